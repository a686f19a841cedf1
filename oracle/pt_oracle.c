/*
 * CPU oracle for the perturbative-triples hot path of libpt.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker.
 * The product library (libpt_amd) never links, loads or calls it.
 *
 * It is a plain-C restatement of the reference algorithm:
 *   src/pt.c:40-185   contraction index maps and disconnected terms
 *   src/pt.c:187-452  (T) same-spin / mixed-spin cases
 *   src/pt.c:537-934  (fT) same-spin / mixed-spin cases
 *   src/pt.c:454-535, 936-1050  block offsets, factor 2, spin-case sums
 *   src/ptmp.c        the float ("_mp") variant of all of the above
 * with its own loop structure (no BLAS).  The energy of every ijk item is
 * accumulated with compensated summation and items are summed in a fixed
 * order, so the result does not depend on the OpenMP thread count.
 *
 * Build: oracle/Makefile (gcc -std=c11, so floating-point contraction is
 * off and every product and sum rounds separately).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static void *
ora_alloc(size_t n)
{
	void *p = malloc(n ? n : 1);

	if (p == NULL) {
		fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
		exit(1);
	}
	return p;
}

/* ijk items in the reference's enumeration order (src/pt.c:210-221 for
 * same-spin: i<j<k; :333-344 for mixed-spin: i<j<o1, any k<o2). */
static int32_t *
ora_ijk_list(size_t o1, size_t o2, int same_spin, size_t n)
{
	int32_t *l = ora_alloc(3 * (n ? n : 1) * sizeof(int32_t));
	size_t i, j, k, it = 0;

	for (i = 0; i < o1; i++)
		for (j = i + 1; j < o1; j++)
			for (k = same_spin ? j + 1 : 0;
			    k < (same_spin ? o1 : o2); k++, it++) {
				l[3 * it] = (int32_t)i;
				l[3 * it + 1] = (int32_t)j;
				l[3 * it + 2] = (int32_t)k;
			}
	return l;
}

/* Sizes of the four spin blocks aaaa/abab/bbbb/baba, src/pt.c:487-512. */
struct ora_blocks {
	size_t ov, oovv[4], oovo[4], ovvv[4];
};

static struct ora_blocks
ora_split(size_t oa, size_t va, size_t ob, size_t vb)
{
	struct ora_blocks s;

	s.ov = oa * va;
	s.oovv[0] = oa * oa * va * va;
	s.oovv[1] = oa * ob * va * vb;
	s.oovv[2] = ob * ob * vb * vb;
	s.oovv[3] = ob * oa * vb * va;
	s.oovo[0] = oa * oa * va * oa;
	s.oovo[1] = oa * ob * va * ob;
	s.oovo[2] = ob * ob * vb * ob;
	s.oovo[3] = ob * oa * vb * oa;
	s.ovvv[0] = oa * va * va * (va - 1) / 2;
	s.ovvv[1] = oa * vb * va * vb;
	s.ovvv[2] = ob * vb * vb * (vb - 1) / 2;
	s.ovvv[3] = ob * va * vb * va;
	return s;
}

#define R double
#define FN(x) x##_d
#include "pt_oracle_body.h"
#undef R
#undef FN

#define R float
#define FN(x) x##_f
#include "pt_oracle_body.h"
#undef R
#undef FN
