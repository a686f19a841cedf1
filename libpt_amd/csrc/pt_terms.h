/* Term tables of the spin cases (pt_terms.cpp). */
#ifndef LIBPT_AMD_PT_TERMS_H
#define LIBPT_AMD_PT_TERMS_H

#include <stdlib.h>

#include "pt_engine.h"

/* A fixed-capacity array with the parts of std::vector the host code uses.
 * It never allocates, so plans built from it (and the allocation-free
 * entry point's plan on the stack) need no heap; growing past N is a
 * programming error and aborts. */
template <class T, int N>
struct FixedVec {
	T v[N];
	int n = 0;
	void push_back(const T &x)
	{
		if (n >= N)
			abort();
		v[n++] = x;
	}
	int size() const { return n; }
	bool empty() const { return n == 0; }
	void clear() { n = 0; }
	T &operator[](int i) { return v[i]; }
	const T &operator[](int i) const { return v[i]; }
	T *data() { return v; }
	const T *data() const { return v; }
	T *begin() { return v; }
	T *end() { return v + n; }
	const T *begin() const { return v; }
	const T *end() const { return v + n; }
};

/* the terms of one spin case, both W sets */
typedef FixedVec<pt_term, PT_MAX_TERMS> pt_term_list;

/* same-spin case of sizes (o, v): 18 terms for W set `set` */
void pt_terms_same(pt_term_list &out, int set, int64_t o, int64_t v);
/* mixed-spin case, pair spin (oa, va), odd spin (ob, vb): 16 terms.
 * vc_t: the VC operand is packed with its plane transposed ([b][c], c
 * contiguous), so that the pairs of one DMA row are contiguous */
void pt_terms_mixed(pt_term_list &out, int set, int64_t oa,
    int64_t va, int64_t ob, int64_t vb, bool vc_t);
/* order by W set, then role a, role b (phase 1), role c (phase 2), stable
 * within each group; returns the number of phase-1 terms of one set,
 * *n_role_a the number of role-a terms of one set */
int pt_terms_order(pt_term_list &terms, int *n_role_a);

#endif
