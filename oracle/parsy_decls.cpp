// TEST INFRASTRUCTURE ONLY.  Takes the address of each supernodal solve that
// the reference's parsy/triangularSolve/Triangular_BCSC.h declares, so that
// the object's undefined symbols are the reference's own mangled names; the
// tests check that the product library exports every one of them.
#include <triangularSolve/Triangular_BCSC.h>

namespace P = sym_lib::parsy;
void *symref_parsy_decls[] = {
    (void *)&P::blockedLsolve,          (void *)&P::blockedLsolve_mrhs,
    (void *)&P::blockedLTsolve,         (void *)&P::blockedLTsolve_mrhs,
    (void *)&P::H2LeveledBlockedLsolve, (void *)&P::H2LeveledBlockedLsolve_mrhs,
    (void *)&P::H2LeveledBlockedLTsolve, (void *)&P::H2LeveledBlockedLTsolve_mrhs,
};
