/*
 * Term tables: every contraction of the connected triples W of libpt's
 * spin cases, written as (role, sign, S slice, Q slice) for the tile
 * kernel.  Index maps restate src/pt.c:40-160 (GEMM wrappers) combined with
 * the permutation passes of src/pt.c:236-290 (same spin) and
 * src/pt.c:361-432 (mixed spin); the (fT) builds of src/pt.c:588-696 and
 * :773-915 reuse the same tables on other operand buffers.
 *
 * Occupied slots: 0 = i, 1 = j, 2 = k.  Operand buffer slots (per W set s,
 * base 12*s): TA TB TC OA OB OC in the API layout (S operands), VA VB VC
 * and TA TB TC packed (Q operands); the same-spin case uses the A slots.
 */
#include <string.h>

#include "pt_engine.h"
#include "pt_terms.h"

namespace {

enum { I = 0, J = 1, K = 2 };
enum { A = PT_ROLE_A, B = PT_ROLE_B, C = PT_ROLE_C };

struct TermBuilder {
	pt_term t;
	TermBuilder(int role, int set, int sign, int kext)
	{
		memset(&t, 0, sizeof(t));
		t.role = role;
		t.set = set;
		t.sign = sign;
		t.K = kext;
	}
	/* S[s][kappa] = buf[sbuf] + sum occ*coef + s*s_st + kappa */
	TermBuilder &S(int buf, int o1, int64_t c1, int o2, int64_t c2,
	    int64_t st)
	{
		t.s_buf = buf;
		t.s_off[o1] += c1;
		t.s_off[o2] += c2;
		t.s_st = st;
		return *this;
	}
	/* Q[kappa][..] = buf[qbuf] + occ*coef + kappa*q_k + strides */
	TermBuilder &Q(int buf, int o1, int64_t c1, int64_t qk)
	{
		t.q_buf = buf;
		t.q_off[o1] += c1;
		t.q_k = qk;
		return *this;
	}
	TermBuilder &st(int role, int64_t stride)
	{
		t.q_st[role] = stride;
		return *this;
	}
	TermBuilder &packed()
	{
		t.packed = 1;
		return *this;
	}
	pt_term done()
	{
		/* the pair (u, w) is the two non-S roles in a, b, c order */
		int u = t.role == A ? B : A;
		int w = t.role == C ? B : C;
		t.ufast = !t.packed && t.q_st[u] == 1 && t.q_st[w] != 1;
		(void)w;
		return t;
	}
};

} /* namespace */

void
pt_terms_same(pt_term_list &out, int set, int64_t o, int64_t v)
{
	const int bT = 12 * set + 0, bO = 12 * set + 3, bV = 12 * set + 6;
	const int bTq = 12 * set + 9;
	const int64_t P = v * (v - 1) / 2, vv = v * v;

	/* ovvv: Y_pqr(x; y,z) = sum_d T[p,q,x,d] V[r,d,(y,z)],
	 * (p,q,r) = (i,j,k) +, (i,k,j) -, (k,j,i) -  (src/pt.c:236-261) */
	const int pqr[3][3] = {{I, J, K}, {I, K, J}, {K, J, I}};
	const int sg1[3] = {+1, -1, -1};
	for (int n = 0; n < 3; ++n) {
		const int p = pqr[n][0], q = pqr[n][1], r = pqr[n][2];
		const int s = sg1[n];
		/* +Y(c; a,b) */
		out.push_back(TermBuilder(C, set, +s, (int)v)
		    .S(bT, p, o * vv, q, vv, v).Q(bV, r, v * P, P).packed().done());
		/* -Y(b; a,c) */
		out.push_back(TermBuilder(B, set, -s, (int)v)
		    .S(bT, p, o * vv, q, vv, v).Q(bV, r, v * P, P).packed().done());
		/* +Y(a; b,c) */
		out.push_back(TermBuilder(A, set, +s, (int)v)
		    .S(bT, p, o * vv, q, vv, v).Q(bV, r, v * P, P).packed().done());
	}
	/* oovo: X_(p;q,r)(x,y,z) = sum_l T[p,l,x,y] O[q,r,z,l],
	 * (p;q,r) = (i;j,k) +, (j;i,k) -, (k;j,i) -  (src/pt.c:263-290) */
	const int pqr2[3][3] = {{I, J, K}, {J, I, K}, {K, J, I}};
	for (int n = 0; n < 3; ++n) {
		const int p = pqr2[n][0], q = pqr2[n][1], r = pqr2[n][2];
		const int s = sg1[n];
		/* +X(b,c,a): S = O[q,r][a][l], Q = T[p][l][b][c] */
		out.push_back(TermBuilder(A, set, +s, (int)o)
		    .S(bO, q, o * v * o, r, v * o, o).Q(bTq, p, o * vv, vv)
		    .st(B, v).st(C, 1).done());
		/* -X(a,c,b): S = O[q,r][b][l], Q = T[p][l][a][c] */
		out.push_back(TermBuilder(B, set, -s, (int)o)
		    .S(bO, q, o * v * o, r, v * o, o).Q(bTq, p, o * vv, vv)
		    .st(A, v).st(C, 1).done());
		/* -X(b,a,c): S = O[q,r][c][l], Q = T[p][l][b][a] */
		out.push_back(TermBuilder(C, set, -s, (int)o)
		    .S(bO, q, o * v * o, r, v * o, o).Q(bTq, p, o * vv, vv)
		    .st(A, 1).st(B, v).done());
	}
}

void
pt_terms_mixed(pt_term_list &out, int set, int64_t oa, int64_t va,
    int64_t ob, int64_t vb, bool vc_t)
{
	const int b0 = 12 * set;
	const int TA = b0 + 0, TB = b0 + 1, TC = b0 + 2;
	const int OA = b0 + 3, OB = b0 + 4, OC = b0 + 5;
	const int VA = b0 + 6, VB = b0 + 7, VC = b0 + 8;
	const int TAq = b0 + 9, TBq = b0 + 10, TCq = b0 + 11;
	const int64_t PA = va * (va - 1) / 2;
	const int ka = (int)va, kb = (int)vb, koa = (int)oa, kob = (int)ob;

	/* src/pt.c:361-367: -sum_d TA[i,j,a,d] VC[k,d,c,b] + (a<->b); the
	 * plane (c, b) of VC, b contiguous, or transposed (b, c) */
	const int64_t vcu = vc_t ? vb : 1, vcw = vc_t ? 1 : va;
	out.push_back(TermBuilder(A, set, -1, ka)
	    .S(TA, I, oa * va * va, J, va * va, va)
	    .Q(VC, K, va * vb * va, vb * va).st(B, vcu).st(C, vcw).done());
	out.push_back(TermBuilder(B, set, +1, ka)
	    .S(TA, I, oa * va * va, J, va * va, va)
	    .Q(VC, K, va * vb * va, vb * va).st(A, vcu).st(C, vcw).done());
	/* :369-375: -sum_d TB[i,k,a,d] VB[j,d,b,c] + (a<->b) */
	out.push_back(TermBuilder(A, set, -1, kb)
	    .S(TB, I, ob * va * vb, K, va * vb, vb)
	    .Q(VB, J, vb * va * vb, va * vb).st(B, vb).st(C, 1).done());
	out.push_back(TermBuilder(B, set, +1, kb)
	    .S(TB, I, ob * va * vb, K, va * vb, vb)
	    .Q(VB, J, vb * va * vb, va * vb).st(A, vb).st(C, 1).done());
	/* :377-383: +sum_d TB[j,k,a,d] VB[i,d,b,c] - (a<->b) */
	out.push_back(TermBuilder(A, set, +1, kb)
	    .S(TB, J, ob * va * vb, K, va * vb, vb)
	    .Q(VB, I, vb * va * vb, va * vb).st(B, vb).st(C, 1).done());
	out.push_back(TermBuilder(B, set, -1, kb)
	    .S(TB, J, ob * va * vb, K, va * vb, vb)
	    .Q(VB, I, vb * va * vb, va * vb).st(A, vb).st(C, 1).done());
	/* :385-392: -sum_d TC[k,j,c,d] VA[i,d,(a,b)] + sum_d TC[k,i,c,d] VA[j,..] */
	out.push_back(TermBuilder(C, set, -1, ka)
	    .S(TC, K, oa * vb * va, J, vb * va, va)
	    .Q(VA, I, va * PA, PA).packed().done());
	out.push_back(TermBuilder(C, set, +1, ka)
	    .S(TC, K, oa * vb * va, I, vb * va, va)
	    .Q(VA, J, va * PA, PA).packed().done());
	/* :394-406: -sum_l TA[i,l,a,b] OC[k,j,c,l] + sum_l TA[j,l,a,b] OC[k,i,c,l] */
	out.push_back(TermBuilder(C, set, -1, koa)
	    .S(OC, K, oa * vb * oa, J, vb * oa, oa)
	    .Q(TAq, I, oa * va * va, va * va).st(A, va).st(B, 1).done());
	out.push_back(TermBuilder(C, set, +1, koa)
	    .S(OC, K, oa * vb * oa, I, vb * oa, oa)
	    .Q(TAq, J, oa * va * va, va * va).st(A, va).st(B, 1).done());
	/* :408-414: -sum_l TB[i,l,a,c] OB[j,k,b,l] + sum_l TB[i,l,b,c] OB[j,k,a,l] */
	out.push_back(TermBuilder(B, set, -1, kob)
	    .S(OB, J, ob * va * ob, K, va * ob, ob)
	    .Q(TBq, I, ob * va * vb, va * vb).st(A, vb).st(C, 1).done());
	out.push_back(TermBuilder(A, set, +1, kob)
	    .S(OB, J, ob * va * ob, K, va * ob, ob)
	    .Q(TBq, I, ob * va * vb, va * vb).st(B, vb).st(C, 1).done());
	/* :416-422: -sum_l TB[j,l,b,c] OB[i,k,a,l] + sum_l TB[j,l,a,c] OB[i,k,b,l] */
	out.push_back(TermBuilder(A, set, -1, kob)
	    .S(OB, I, ob * va * ob, K, va * ob, ob)
	    .Q(TBq, J, ob * va * vb, va * vb).st(B, vb).st(C, 1).done());
	out.push_back(TermBuilder(B, set, +1, kob)
	    .S(OB, I, ob * va * ob, K, va * ob, ob)
	    .Q(TBq, J, ob * va * vb, va * vb).st(A, vb).st(C, 1).done());
	/* :424-432: -sum_l TC[k,l,c,a] OA[j,i,b,l] + sum_l TC[k,l,c,b] OA[j,i,a,l] */
	out.push_back(TermBuilder(B, set, -1, koa)
	    .S(OA, J, oa * va * oa, I, va * oa, oa)
	    .Q(TCq, K, oa * vb * va, vb * va).st(A, 1).st(C, va).done());
	out.push_back(TermBuilder(A, set, +1, koa)
	    .S(OA, J, oa * va * oa, I, va * oa, oa)
	    .Q(TCq, K, oa * vb * va, vb * va).st(B, 1).st(C, va).done());
}

int
pt_terms_order(pt_term_list &terms, int *n_role_a)
{
	/* W set 0, then W set 1; within a set role a, role b, role c, each
	 * group in table order.  Both sets have the same structure, so the
	 * counts of set 0 hold for every set. */
	int nset = 1;
	for (const pt_term &t : terms)
		nset = t.set + 1 > nset ? t.set + 1 : nset;
	pt_term_list out;
	int na = 0, n1 = 0;
	for (int s = 0; s < nset; ++s)
		for (int role = PT_ROLE_A; role <= PT_ROLE_C; ++role)
			for (const pt_term &t : terms)
				if (t.set == s && t.role == role) {
					out.push_back(t);
					if (s == 0 && role == PT_ROLE_A)
						++na;
					if (s == 0 && role != PT_ROLE_C)
						++n1;
				}
	terms = out;
	if (n_role_a)
		*n_role_a = na;
	return n1;
}
