/* Term tables of the spin cases (pt_terms.cpp). */
#ifndef LIBPT_AMD_PT_TERMS_H
#define LIBPT_AMD_PT_TERMS_H

#include <vector>

#include "pt_engine.h"

/* same-spin case of sizes (o, v): 18 terms for W set `set` */
void pt_terms_same(std::vector<pt_term> &out, int set, int64_t o, int64_t v);
/* mixed-spin case, pair spin (oa, va), odd spin (ob, vb): 16 terms */
void pt_terms_mixed(std::vector<pt_term> &out, int set, int64_t oa,
    int64_t va, int64_t ob, int64_t vb);
/* order by W set, then role a, role b (phase 1), role c (phase 2), stable
 * within each group; returns the number of phase-1 terms of one set,
 * *n_role_a the number of role-a terms of one set */
int pt_terms_order(std::vector<pt_term> &terms, int *n_role_a);

#endif
