/*
 * Deterministic synthetic inputs for the perturbative-triples API.
 *
 * Every element of every pt.h input array is a pure function
 *     value = f(seed, tensor, flat index in the API array)
 * built from a counter-based hash, so the same bits come out on the host
 * (gcc, ISO C mode: no floating-point contraction) and on the device
 * (hipcc, contraction disabled inside these functions), for any thread
 * count and any fill order.  Only IEEE +, -, * and sqrt (for index
 * decoding, then corrected in integers) are used.
 *
 * Distributions:
 *   e_occ ~ U[-1.5,-0.3], e_vir ~ U[0.1,3.0], d_ov[i,a] = e_i - e_a + w/3
 *   N(0,s) is approximated by s * sqrt(3) * (u1+u2+u3+u4 - 2) with four
 *   32-bit uniforms (Irwin-Hall, unit variance, |z| <= 2 sqrt 3); the sum
 *   of the uniforms is exact in double.
 *   Restricted blocks obey the closed-shell relations of libpt's inputs
 *   (as in its rpt fixtures):
 *     t2_abab[ijab] = (X[ijab] + X[jiba]) / sqrt 2,  t2_aaaa = abab - abab(a<->b)
 *     oovv likewise; oovo_aaaa[jkcl] = abab[jkcl] - abab[kjcl];
 *     ovvv_aaaa[kd(b>c)] = abab[kdbc] - abab[kdcb]
 *   Unrestricted same-spin blocks are antisymmetrised draws, t2/oovv baba
 *   blocks are the (ji,ba) transpose of abab, oovo/ovvv baba blocks are
 *   independent draws.
 */
#ifndef LIBPT_AMD_SYNTH_H
#define LIBPT_AMD_SYNTH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define PT_HD __host__ __device__ static inline
#else
#define PT_HD static inline
#endif

/* no fused multiply-add anywhere in the generator (clang honours this on
 * host and device; gcc in ISO C mode never contracts) */
#if defined(__clang__)
#define PT_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define PT_NOCONTRACT
#endif

/* tensor ids (the API argument a value belongs to) */
enum pt_synth_tensor {
	PTS_D_OV = 0, PTS_F_OV, PTS_T1, PTS_T2, PTS_OOVO, PTS_OOVV, PTS_OVVV,
	PTS_L1, PTS_L2, PTS_I2, PTS_I3, PTS_I6, PTS_I7, PTS_NTENSOR
};

struct pt_synth_spec {
	uint64_t seed;
	int64_t oa, va, ob, vb;
	int32_t restricted;  /* 1: rpt/rft layout (aaaa|abab), 0: four blocks */
	int32_t pad_;
	double omega;        /* folded into d_ov as omega/3 */
	double sig_t1, sig_t2, sig_int, sig_f;
};

PT_HD uint64_t
pt_mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

PT_HD uint64_t
pt_hash(uint64_t seed, uint32_t stream, uint64_t idx)
{
	uint64_t key = pt_mix64(seed + 0x9e3779b97f4a7c15ULL *
	    (uint64_t)(stream + 1));
	return pt_mix64(key ^ pt_mix64(idx + 0x632be59bd9b4e019ULL));
}

/* uniform in [0,1): 53 random bits, exact */
PT_HD double
pt_u01(uint64_t seed, uint32_t stream, uint64_t idx)
{
	PT_NOCONTRACT
	return (double)(pt_hash(seed, stream, idx) >> 11) *
	    (1.0 / 9007199254740992.0);
}

/* approximately N(0,1): Irwin-Hall of four 32-bit uniforms */
PT_HD double
pt_nrm(uint64_t seed, uint32_t stream, uint64_t idx)
{
	PT_NOCONTRACT
	uint64_t h1 = pt_hash(seed, stream, 2 * idx);
	uint64_t h2 = pt_hash(seed, stream, 2 * idx + 1);
	const double s32 = 1.0 / 4294967296.0;
	double u = (double)(h1 & 0xffffffffULL) * s32;
	u += (double)(h1 >> 32) * s32;
	u += (double)(h2 & 0xffffffffULL) * s32;
	u += (double)(h2 >> 32) * s32;
	return (u - 2.0) * 1.7320508075688772;
}

/* stream ids: family * 8 + block (0 aa, 1 ab, 2 bb, 3 ba) */
enum {
	PTS_S_EOCC = 8, PTS_S_EVIR = 16, PTS_S_F = 24, PTS_S_T1 = 32,
	PTS_S_T2 = 40, PTS_S_OOVV = 48, PTS_S_OOVO = 56, PTS_S_OVVV = 64,
	PTS_S_L1 = 72, PTS_S_L2 = 80, PTS_S_I2 = 88, PTS_S_I3 = 96,
	PTS_S_I6 = 104, PTS_S_I7 = 112, PTS_S_F2 = 120
};

/* b > c from packed index p = b(b-1)/2 + c */
PT_HD void
pt_untri(int64_t p, int64_t *b, int64_t *c)
{
	int64_t bb = (int64_t)((1.0 + __builtin_sqrt(1.0 + 8.0 * (double)p)) *
	    0.5);
	while (bb * (bb - 1) / 2 > p)
		bb--;
	while ((bb + 1) * bb / 2 <= p)
		bb++;
	*b = bb;
	*c = p - bb * (bb - 1) / 2;
}

/* sizes of the four spin blocks of a 4-index family */
PT_HD void
pt_blk4(const struct pt_synth_spec *s, int fam, int64_t n[4])
{
	int64_t oa = s->oa, va = s->va, ob = s->ob, vb = s->vb;
	if (fam == 0) {        /* oovv-like [o][o][v][v] */
		n[0] = oa * oa * va * va; n[1] = oa * ob * va * vb;
		n[2] = ob * ob * vb * vb; n[3] = ob * oa * vb * va;
	} else if (fam == 1) { /* oovo [o][o][v][o] */
		n[0] = oa * oa * va * oa; n[1] = oa * ob * va * ob;
		n[2] = ob * ob * vb * ob; n[3] = ob * oa * vb * oa;
	} else {               /* ovvv [o][v][v][v], same-spin packed */
		n[0] = oa * va * (va * (va - 1) / 2); n[1] = oa * vb * va * vb;
		n[2] = ob * vb * (vb * (vb - 1) / 2); n[3] = ob * va * vb * va;
	}
}

/* one amplitude-like [o1][o2][v1][v2] draw on a dense grid */
PT_HD double
pt_x4(const struct pt_synth_spec *s, uint32_t st, double sig, int64_t o2,
    int64_t v1, int64_t v2, int64_t i, int64_t j, int64_t a, int64_t b)
{
	PT_NOCONTRACT
	return sig * pt_nrm(s->seed, st, ((i * o2 + j) * v1 + a) * v2 + b);
}

/* oovv-like element: family stream base st, block blk, indices */
PT_HD double
pt_oovv_elem(const struct pt_synth_spec *s, uint32_t st, double sig, int blk,
    int64_t i, int64_t j, int64_t a, int64_t b)
{
	PT_NOCONTRACT
	int64_t oa = s->oa, va = s->va, ob = s->ob, vb = s->vb;
	if (s->restricted) {
		/* abab(i,j,a,b) = (X[ijab] + X[jiba]) / sqrt 2 */
		double x, y, abab_ab, abab_ba;
		x = pt_x4(s, st + 1, sig, oa, va, va, i, j, a, b);
		y = pt_x4(s, st + 1, sig, oa, va, va, j, i, b, a);
		abab_ab = (x + y) * 0.7071067811865476;
		if (blk == 1)
			return abab_ab;
		x = pt_x4(s, st + 1, sig, oa, va, va, i, j, b, a);
		y = pt_x4(s, st + 1, sig, oa, va, va, j, i, a, b);
		abab_ba = (x + y) * 0.7071067811865476;
		return abab_ab - abab_ba;
	}
	if (blk == 0 || blk == 2) {
		int64_t o = blk == 0 ? oa : ob, v = blk == 0 ? va : vb;
		uint32_t ss = st + (uint32_t)blk;
		double p = pt_x4(s, ss, sig, o, v, v, i, j, a, b) -
		    pt_x4(s, ss, sig, o, v, v, i, j, b, a);
		double q = pt_x4(s, ss, sig, o, v, v, j, i, a, b) -
		    pt_x4(s, ss, sig, o, v, v, j, i, b, a);
		return (p - q) * 0.5;
	}
	if (blk == 1)
		return pt_x4(s, st + 1, sig, ob, va, vb, i, j, a, b);
	/* baba(i,j,a,b) = abab(j,i,b,a) */
	return pt_x4(s, st + 1, sig, ob, va, vb, j, i, b, a);
}

PT_HD double
pt_oovo_elem(const struct pt_synth_spec *s, uint32_t st, double sig, int blk,
    int64_t j, int64_t k, int64_t c, int64_t l)
{
	PT_NOCONTRACT
	int64_t oa = s->oa, va = s->va, ob = s->ob, vb = s->vb;
	if (s->restricted) {
		double x = pt_x4(s, st + 1, sig, oa, va, oa, j, k, c, l);
		if (blk == 1)
			return x;
		return x - pt_x4(s, st + 1, sig, oa, va, oa, k, j, c, l);
	}
	if (blk == 0 || blk == 2) {
		int64_t o = blk == 0 ? oa : ob, v = blk == 0 ? va : vb;
		uint32_t ss = st + (uint32_t)blk;
		return pt_x4(s, ss, sig, o, v, o, j, k, c, l) -
		    pt_x4(s, ss, sig, o, v, o, k, j, c, l);
	}
	if (blk == 1)
		return pt_x4(s, st + 1, sig, ob, va, ob, j, k, c, l);
	return pt_x4(s, st + 3, sig, oa, vb, oa, j, k, c, l);
}

PT_HD double
pt_ovvv_elem(const struct pt_synth_spec *s, uint32_t st, double sig, int blk,
    int64_t k, int64_t d, int64_t b, int64_t c)
{
	PT_NOCONTRACT
	int64_t va = s->va, vb = s->vb;
	if (s->restricted) {
		double x = pt_x4(s, st + 1, sig, va, va, va, k, d, b, c);
		if (blk == 1)
			return x;
		return x - pt_x4(s, st + 1, sig, va, va, va, k, d, c, b);
	}
	if (blk == 0 || blk == 2) {
		int64_t v = blk == 0 ? va : vb;
		uint32_t ss = st + (uint32_t)blk;
		return pt_x4(s, ss, sig, v, v, v, k, d, b, c) -
		    pt_x4(s, ss, sig, v, v, v, k, d, c, b);
	}
	if (blk == 1)
		return pt_x4(s, st + 1, sig, vb, va, vb, k, d, b, c);
	return pt_x4(s, st + 3, sig, va, vb, va, k, d, b, c);
}

/* Element `idx` of API array `t` (flat index over all its blocks). */
PT_HD double
pt_synth_elem(const struct pt_synth_spec *s, int t, int64_t idx)
{
	PT_NOCONTRACT
	int64_t oa = s->oa, va = s->va, ob = s->ob, vb = s->vb;
	int64_t n[4], o1, o2, v1, v2, i, j, a, b;
	int blk, fam, spin;
	uint32_t st = 0;
	double sig = 0;

	if (t == PTS_D_OV || t == PTS_F_OV || t == PTS_T1 || t == PTS_L1) {
		spin = idx >= oa * va;
		if (spin)
			idx -= oa * va;
		v1 = spin ? vb : va;
		i = idx / v1;
		a = idx % v1;
		if (t == PTS_D_OV) {
			double eo = -1.5 + 1.2 * pt_u01(s->seed,
			    PTS_S_EOCC + (uint32_t)spin, (uint64_t)i);
			double ev = 0.1 + 2.9 * pt_u01(s->seed,
			    PTS_S_EVIR + (uint32_t)spin, (uint64_t)a);
			return (eo - ev) + s->omega / 3.0;
		}
		st = t == PTS_F_OV ? PTS_S_F : t == PTS_T1 ? PTS_S_T1 : PTS_S_L1;
		sig = t == PTS_F_OV ? s->sig_f : s->sig_t1;
		return sig * pt_nrm(s->seed, st + (uint32_t)spin, (uint64_t)idx);
	}
	switch (t) {
	case PTS_T2: fam = 0; st = PTS_S_T2; sig = s->sig_t2; break;
	case PTS_L2: fam = 0; st = PTS_S_L2; sig = s->sig_t2; break;
	case PTS_OOVV: fam = 0; st = PTS_S_OOVV; sig = s->sig_int; break;
	case PTS_OOVO: fam = 1; st = PTS_S_OOVO; sig = s->sig_int; break;
	case PTS_I2: fam = 1; st = PTS_S_I2; sig = s->sig_int; break;
	case PTS_I6: fam = 1; st = PTS_S_I6; sig = s->sig_int; break;
	case PTS_OVVV: fam = 2; st = PTS_S_OVVV; sig = s->sig_int; break;
	case PTS_I3: fam = 2; st = PTS_S_I3; sig = s->sig_int; break;
	default: fam = 2; st = PTS_S_I7; sig = s->sig_int; break;
	}
	pt_blk4(s, fam, n);
	for (blk = 0; blk < 3 && idx >= n[blk]; blk++)
		idx -= n[blk];
	/* block geometry: (o1, o2, v1, v2) of [p][q][x][y] */
	if (blk == 0) { o1 = oa; o2 = oa; v1 = va; v2 = va; }
	else if (blk == 1) { o1 = oa; o2 = ob; v1 = va; v2 = vb; }
	else if (blk == 2) { o1 = ob; o2 = ob; v1 = vb; v2 = vb; }
	else { o1 = ob; o2 = oa; v1 = vb; v2 = va; }
	(void)o1;
	if (fam == 0) {
		b = idx % v2; idx /= v2;
		a = idx % v1; idx /= v1;
		j = idx % o2; i = idx / o2;
		return pt_oovv_elem(s, st, sig, blk, i, j, a, b);
	}
	if (fam == 1) {
		/* [j][k][c][l]: k, l in o2 (the second occupied spin) */
		int64_t l = idx % o2; idx /= o2;
		a = idx % v1; idx /= v1;
		j = idx % o2; i = idx / o2;
		return pt_oovo_elem(s, st, sig, blk, i, j, a, l);
	}
	/* ovvv [k][d][b][c]: same-spin packed (b>c); mixed [k][d:v2][b:v1][c:v2] */
	if (blk == 0 || blk == 2) {
		int64_t v = blk == 0 ? va : vb, P = v * (v - 1) / 2, p;
		p = idx % P; idx /= P;
		a = idx % v; i = idx / v;
		pt_untri(p, &b, &j);
		return pt_ovvv_elem(s, st, sig, blk, i, a, b, j);
	}
	{
		int64_t c = idx % v2; idx /= v2;
		b = idx % v1; idx /= v1;
		a = idx % v2; i = idx / v2;
		return pt_ovvv_elem(s, st, sig, blk, i, a, b, c);
	}
}

/* number of elements of API array t */
PT_HD int64_t
pt_synth_size(const struct pt_synth_spec *s, int t)
{
	int64_t n[4];
	int fam;

	if (t == PTS_D_OV || t == PTS_F_OV || t == PTS_T1 || t == PTS_L1)
		return s->oa * s->va + (s->restricted ? 0 : s->ob * s->vb);
	fam = (t == PTS_T2 || t == PTS_L2 || t == PTS_OOVV) ? 0 :
	    (t == PTS_OOVO || t == PTS_I2 || t == PTS_I6) ? 1 : 2;
	pt_blk4(s, fam, n);
	return s->restricted ? n[0] + n[1] : n[0] + n[1] + n[2] + n[3];
}

#endif /* LIBPT_AMD_SYNTH_H */
