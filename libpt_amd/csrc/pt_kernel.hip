/*
 * CDNA4 (gfx950) kernels of the perturbative-triples engine.
 *
 * k_tiles: one workgroup = one ijk item x one (A, B, C) tile of 32 a x 32 b
 * x TCW c (TCW = 32 for (T), 16 for (fT)).  Four waves each own 16 x 16 x TCW
 * (a, b, c) block of the connected triples W in MFMA accumulators and never
 * write W to memory:
 *
 *   phase 1 -- accumulator layout [a][b]@c (one 16x16 MFMA block per c):
 *       terms whose single index is a  (A = S rows a,  B = Q[.][b][c])
 *       terms whose single index is b  (A = Q[.][a][c], B = S rows b)
 *   transpose through LDS to layout [a][c]@b (one block per b)
 *   phase 2 -- terms whose single index is c
 *       (A = Q[.][a][b], B = S rows c)
 *   epilogue -- for every (a, b, c) of the reference's triangle:
 *       disconnected term Z, denominator Dn, energy W (W - Z) / Dn
 *       ((fT): (W_L - Z_L) W_R / Dn), compensated FP64 sum per lane,
 *       one deterministic reduction per workgroup into partial[item].
 *
 * Operand tiles are staged global -> LDS per kappa chunk with the
 * contiguous global index mapped to consecutive threads.
 *
 * Forms: FP64 (v_mfma_f64_16x16x4_f64), FP32 (v_mfma_f32_16x16x4_f32),
 * split BF16 / FP16 (hi+lo pairs, hi*hi + hi*lo + lo*hi with
 * v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulation, FP64 epilogue).
 *
 * MFMA lane layouts (checked on gfx950): A[m][k] lane m + 16*(k/KL),
 * B[k][n] lane n + 16*(k/KL) (KL = 1 for x4, 8 consecutive k for x32);
 * D[m][n] lane n + 16*(m/4) reg m%4 for f32 results, but
 * lane n + 16*(m%4) reg m/4 for v_mfma_f64_16x16x4_f64.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_engine.h"
#include "synth.h"

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef uint16_t v8u __attribute__((ext_vector_type(8)));

namespace {

template <int F> struct Form;

template <> struct Form<PT_FORM_FP64> {
	typedef double G;   /* operand element in HBM */
	typedef double L;   /* LDS element (one plane) */
	typedef double AE;  /* accumulator element */
	typedef v4d ACC;
	typedef double ER;  /* epilogue real */
	static constexpr int NP = 1, KC = 8, KS = 4;
	static constexpr bool ROW_F64 = true, SPLIT = false;
};

template <> struct Form<PT_FORM_FP32> {
	typedef float G;
	typedef float L;
	typedef float AE;
	typedef v4f ACC;
	typedef float ER;
	static constexpr int NP = 1, KC = 16, KS = 4;
	static constexpr bool ROW_F64 = false, SPLIT = false;
};

template <> struct Form<PT_FORM_BF16X3> {
	typedef uint32_t G;  /* hi in bits 0-15, lo in bits 16-31 */
	typedef uint16_t L;
	typedef float AE;
	typedef v4f ACC;
	typedef double ER;
	static constexpr int NP = 2, KC = 32, KS = 32;
	static constexpr bool ROW_F64 = false, SPLIT = true;
};

template <> struct Form<PT_FORM_F16X3> : Form<PT_FORM_BF16X3> {};

/*
 * LDS: S rows [32][KCP] and Q rows [32 u][QP w][KCP] per plane.  Q rows are
 * skewed (QP = 33) so 16 consecutive u at one w fall in different banks;
 * FP64/FP32 rows are padded by one element, split rows (64 bytes of
 * halves) have their 16-byte chunks XOR-swizzled instead.
 */
template <int F> struct Lds {
	typedef Form<F> FM;
	static constexpr int KCP = FM::SPLIT ? FM::KC : FM::KC + 1;
	static constexpr int QP = 33;
	static constexpr int S_PLANE = 32 * KCP;
	static constexpr int Q_PLANE = 32 * QP * KCP;
	static constexpr int STAGE_BYTES =
	    FM::NP * (S_PLANE + Q_PLANE) * (int)sizeof(typename FM::L);
	static constexpr int X_BYTES = 4 * 32 * 64 * (int)sizeof(typename FM::AE);
	static constexpr int BYTES =
	    STAGE_BYTES > X_BYTES ? STAGE_BYTES : X_BYTES;
	static __device__ __forceinline__ int off(int row, int k)
	{
		if constexpr (FM::SPLIT)
			return row * 32 + (((k >> 3) ^ ((row >> 2) & 3)) << 3) + (k & 7);
		else
			return row * KCP + k;
	}
};

/* ---------------------------------------------------------------- MFMA */

__device__ __forceinline__ v4d
mma(double a, double b, v4d c)
{
	return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4f
mma(float a, float b, v4f c)
{
	return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int F>
__device__ __forceinline__ v4f
mma8(v8u a, v8u b, v4f c)
{
	if constexpr (F == PT_FORM_BF16X3)
		return __builtin_amdgcn_mfma_f32_16x16x32_bf16(
		    __builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c,
		    0, 0, 0);
	else
		return __builtin_amdgcn_mfma_f32_16x16x32_f16(
		    __builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c,
		    0, 0, 0);
}

template <class G>
__device__ __forceinline__ G
neg_if(G v, int neg)
{
	if constexpr (sizeof(G) == 4 && !__is_same(G, float))
		return neg ? (v ^ 0x80008000u) : v; /* both halves */
	else
		return neg ? -v : v;
}

/* ---------------------------------------------------------- epilogue */

template <class ER>
__device__ __forceinline__ ER
z_same(const ER *u, const ER *g, int64_t o, int64_t v, int i, int j, int k,
    int a, int b, int c)
{
	/* src/pt.c:162-175 */
	const int64_t vv = v * v;
	const ER *gjk = g + ((int64_t)j * o + k) * vv;
	const ER *gik = g + ((int64_t)i * o + k) * vv;
	const ER *gji = g + ((int64_t)j * o + i) * vv;
	return +u[i * v + a] * gjk[b * v + c]
	       -u[j * v + a] * gik[b * v + c]
	       -u[k * v + a] * gji[b * v + c]
	       -u[i * v + b] * gjk[a * v + c]
	       +u[j * v + b] * gik[a * v + c]
	       +u[k * v + b] * gji[a * v + c]
	       -u[i * v + c] * gjk[b * v + a]
	       +u[j * v + c] * gik[b * v + a]
	       +u[k * v + c] * gji[b * v + a];
}

template <class ER>
__device__ __forceinline__ ER
z_mixed(const pt_epi &e, int64_t oa, int64_t va, int64_t ob, int64_t vb,
    int i, int j, int k, int a, int b, int c)
{
	/* src/pt.c:177-185 and :433-442 */
	const ER *u1 = (const ER *)e.u1, *u2 = (const ER *)e.u2;
	const ER *f1 = (const ER *)e.f1, *f2 = (const ER *)e.f2;
	const ER *GA = (const ER *)e.GA, *GB = (const ER *)e.GB;
	const ER *ZA = (const ER *)e.ZA, *ZB = (const ER *)e.ZB;
	const int64_t jk = ((int64_t)j * ob + k) * va, ik = ((int64_t)i * ob + k) * va;
	const int64_t jkbc = (jk + b) * vb + c, jkac = (jk + a) * vb + c;
	const int64_t ikbc = (ik + b) * vb + c, ikac = (ik + a) * vb + c;
	const int64_t jiba = (((int64_t)j * oa + i) * va + b) * va + a;
	return (u1[i * va + a] * GB[jkbc] + f1[i * va + a] * ZB[jkbc])
	     - (u1[i * va + b] * GB[jkac] + f1[i * va + b] * ZB[jkac])
	     - (u1[j * va + a] * GB[ikbc] + f1[j * va + a] * ZB[ikbc])
	     + (u1[j * va + b] * GB[ikac] + f1[j * va + b] * ZB[ikac])
	     + (u2[k * vb + c] * GA[jiba] + f2[k * vb + c] * ZA[jiba]);
}

__device__ __forceinline__ void
kahan(double &s, double &cmp, double x)
{
	double y = x - cmp;
	double t = s + y;
	cmp = (t - s) - y;
	s = t;
}

/* ------------------------------------------------- one staging step */

/* Operand geometry of one term inside the workgroup tile (uniform). */
struct Step {
	const void *S, *Q;
	int64_t s_st, q_k, q_u, q_w;
	int role, set, K, neg, packed, ufast;
	int s0, sn, sext, u0, uext, w0, wext, qwe, qsh;
};

__device__ __forceinline__ Step
make_step(const pt_launch &p, const pt_term &T, int i, int j, int k, int A0,
    int B0, int C0, int tcw)
{
	Step s;
	s.S = (const char *)p.buf[T.s_buf];
	s.Q = (const char *)p.buf[T.q_buf];
	s.role = T.role;
	s.set = T.set;
	s.K = T.K;
	s.neg = T.sign < 0;
	s.packed = T.packed;
	s.ufast = T.ufast;
	s.s_st = T.s_st;
	s.q_k = T.q_k;
	const int na = p.na, nc = p.nc;
	int ur, wr;
	if (T.role == PT_ROLE_A) {
		s.s0 = A0; s.sn = 32; s.sext = na;
		s.u0 = B0; s.uext = na; s.w0 = C0; s.wext = nc; s.qwe = tcw;
		ur = PT_ROLE_B; wr = PT_ROLE_C;
	} else if (T.role == PT_ROLE_B) {
		s.s0 = B0; s.sn = 32; s.sext = na;
		s.u0 = A0; s.uext = na; s.w0 = C0; s.wext = nc; s.qwe = tcw;
		ur = PT_ROLE_A; wr = PT_ROLE_C;
	} else {
		s.s0 = C0; s.sn = tcw; s.sext = nc;
		s.u0 = A0; s.uext = na; s.w0 = B0; s.wext = na; s.qwe = 32;
		ur = PT_ROLE_A; wr = PT_ROLE_B;
	}
	s.qsh = s.qwe == 32 ? 5 : 4;
	s.q_u = T.q_st[ur];
	s.q_w = T.q_st[wr];
	s.S = (const char *)s.S; /* offsets applied by the caller's type */
	return s;
}

/* pairs (u, w) of this thread for a step: pair index tid + 256 m */
__device__ __forceinline__ void
pair_uw(const Step &s, int pidx, int &u, int &w)
{
	if (s.ufast) {
		u = pidx & 31;
		w = pidx >> 5;
	} else {
		w = pidx & (s.qwe - 1);
		u = pidx >> s.qsh;
	}
}

template <int F, int NPR, int SEPT>
__device__ __forceinline__ void
load_step(const Step &s, int64_t soff, int64_t qoff, int kc, int tid,
    typename Form<F>::G (&qr)[NPR][Form<F>::KC],
    typename Form<F>::G (&sr)[SEPT])
{
	typedef typename Form<F>::G G;
	constexpr int KC = Form<F>::KC;
	const G *S = (const G *)s.S + soff;
	const G *Q = (const G *)s.Q + qoff;
	/* S: thread owns SEPT consecutive kappa of one row */
	{
		const int row = tid / (KC / SEPT), k0 = (tid % (KC / SEPT)) * SEPT;
		const int sg = s.s0 + row;
		const bool rv = row < s.sn && sg < s.sext;
		const G *sp = S + (int64_t)(rv ? sg : 0) * s.s_st;
#pragma unroll
		for (int e = 0; e < SEPT; ++e) {
			const int kg = kc + k0 + e;
			sr[e] = (rv && kg < s.K) ? neg_if(sp[kg], s.neg) : G(0);
		}
	}
	const int npr = (32 * s.qwe) >> 8;
#pragma unroll
	for (int m = 0; m < NPR; ++m) {
		if (m >= npr)
			break;
		int u, w;
		pair_uw(s, tid + 256 * m, u, w);
		const int ug = s.u0 + u, wg = s.w0 + w;
		const bool v = ug < s.uext && wg < s.wext && (!s.packed || ug > wg);
		int64_t base = 0;
		if (v)
			base = s.packed ? (int64_t)ug * (ug - 1) / 2 + wg :
			    (int64_t)ug * s.q_u + (int64_t)wg * s.q_w;
		const G *qp = Q + base + (int64_t)kc * s.q_k;
		const int kmax = v ? s.K - kc : 0;
#pragma unroll
		for (int kk = 0; kk < KC; ++kk)
			qr[m][kk] = kk < kmax ? qp[(int64_t)kk * s.q_k] : G(0);
	}
}

template <int F, int NPR, int SEPT>
__device__ __forceinline__ void
store_step(const Step &s, int tid, typename Form<F>::L *sS,
    typename Form<F>::L *sQ, const typename Form<F>::G (&qr)[NPR][Form<F>::KC],
    const typename Form<F>::G (&sr)[SEPT])
{
	typedef Form<F> FM;
	typedef Lds<F> LD;
	constexpr int KC = FM::KC;
	{
		const int row = tid / (KC / SEPT), k0 = (tid % (KC / SEPT)) * SEPT;
#pragma unroll
		for (int e = 0; e < SEPT; ++e) {
			const int o = LD::off(row, k0 + e);
			if constexpr (FM::SPLIT) {
				sS[o] = (uint16_t)(sr[e] & 0xffffu);
				sS[LD::S_PLANE + o] = (uint16_t)(sr[e] >> 16);
			} else {
				sS[o] = sr[e];
			}
		}
	}
	const int npr = (32 * s.qwe) >> 8;
#pragma unroll
	for (int m = 0; m < NPR; ++m) {
		if (m >= npr)
			break;
		int u, w;
		pair_uw(s, tid + 256 * m, u, w);
		const int row = u * LD::QP + w;
		if constexpr (FM::SPLIT) {
#pragma unroll
			for (int q = 0; q < 4; ++q) {
				v8u h, l;
#pragma unroll
				for (int e = 0; e < 8; ++e) {
					h[e] = (uint16_t)(qr[m][8 * q + e] & 0xffffu);
					l[e] = (uint16_t)(qr[m][8 * q + e] >> 16);
				}
				const int o = LD::off(row, 8 * q);
				*(v8u *)(sQ + o) = h;
				*(v8u *)(sQ + LD::Q_PLANE + o) = l;
			}
		} else {
#pragma unroll
			for (int kk = 0; kk < KC; ++kk)
				sQ[LD::off(row, kk)] = qr[m][kk];
		}
	}
}

/* MFMA over one staged chunk into a fixed accumulator set */
template <int F, int NCB>
__device__ __forceinline__ void
mma_step(typename Form<F>::ACC (&ac)[16 * NCB], int role,
    const typename Form<F>::L *sS, const typename Form<F>::L *sQ, int wa,
    int wb, int ln, int lg)
{
	typedef Form<F> FM;
	typedef Lds<F> LD;
	typedef typename FM::L L;
	constexpr int KC = FM::KC, KS = FM::KS, QP = LD::QP;
	constexpr int TCW = 16 * NCB;
	if constexpr (!FM::SPLIT) {
#pragma unroll
		for (int ks = 0; ks < KC; ks += KS) {
			const int kap = ks + lg;
			if (role == PT_ROLE_A) {
				const L av = sS[LD::off(16 * wa + ln, kap)];
				const int r0 = (16 * wb + ln) * QP;
#pragma unroll
				for (int x = 0; x < TCW; ++x)
					ac[x] = mma(av, sQ[LD::off(r0 + x, kap)], ac[x]);
			} else if (role == PT_ROLE_B) {
				const L bv = sS[LD::off(16 * wb + ln, kap)];
				const int r0 = (16 * wa + ln) * QP;
#pragma unroll
				for (int x = 0; x < TCW; ++x)
					ac[x] = mma(sQ[LD::off(r0 + x, kap)], bv, ac[x]);
			} else {
				L bv[NCB];
#pragma unroll
				for (int cb = 0; cb < NCB; ++cb)
					bv[cb] = sS[LD::off(16 * cb + ln, kap)];
				const int r0 = (16 * wa + ln) * QP + 16 * wb;
#pragma unroll
				for (int b = 0; b < 16; ++b) {
					const L av = sQ[LD::off(r0 + b, kap)];
#pragma unroll
					for (int cb = 0; cb < NCB; ++cb)
						ac[16 * cb + b] = mma(av, bv[cb], ac[16 * cb + b]);
				}
			}
		}
	} else {
#define LD8(base, plane, row) \
	(*(const v8u *)((base) + (plane) + (row) * 32 + \
	    ((lg ^ (((row) >> 2) & 3)) << 3)))
		if (role == PT_ROLE_A) {
			const int rs = 16 * wa + ln;
			const v8u ah = LD8(sS, 0, rs), al = LD8(sS, LD::S_PLANE, rs);
			const int r0 = (16 * wb + ln) * QP;
#pragma unroll
			for (int x = 0; x < TCW; ++x) {
				const v8u bh = LD8(sQ, 0, r0 + x);
				const v8u bl = LD8(sQ, LD::Q_PLANE, r0 + x);
				ac[x] = mma8<F>(al, bh, ac[x]);
				ac[x] = mma8<F>(ah, bl, ac[x]);
				ac[x] = mma8<F>(ah, bh, ac[x]);
			}
		} else if (role == PT_ROLE_B) {
			const int rs = 16 * wb + ln;
			const v8u bh = LD8(sS, 0, rs), bl = LD8(sS, LD::S_PLANE, rs);
			const int r0 = (16 * wa + ln) * QP;
#pragma unroll
			for (int x = 0; x < TCW; ++x) {
				const v8u ah = LD8(sQ, 0, r0 + x);
				const v8u al = LD8(sQ, LD::Q_PLANE, r0 + x);
				ac[x] = mma8<F>(al, bh, ac[x]);
				ac[x] = mma8<F>(ah, bl, ac[x]);
				ac[x] = mma8<F>(ah, bh, ac[x]);
			}
		} else {
			v8u bh[NCB], bl[NCB];
#pragma unroll
			for (int cb = 0; cb < NCB; ++cb) {
				bh[cb] = LD8(sS, 0, 16 * cb + ln);
				bl[cb] = LD8(sS, LD::S_PLANE, 16 * cb + ln);
			}
			const int r0 = (16 * wa + ln) * QP + 16 * wb;
#pragma unroll
			for (int b = 0; b < 16; ++b) {
				const v8u ah = LD8(sQ, 0, r0 + b);
				const v8u al = LD8(sQ, LD::Q_PLANE, r0 + b);
#pragma unroll
				for (int cb = 0; cb < NCB; ++cb) {
					const int x = 16 * cb + b;
					ac[x] = mma8<F>(al, bh[cb], ac[x]);
					ac[x] = mma8<F>(ah, bl[cb], ac[x]);
					ac[x] = mma8<F>(ah, bh[cb], ac[x]);
				}
			}
		}
#undef LD8
	}
}

/* -------------------------------------------------------------- kernel */

template <int F, int KIND, int NSET>
__global__ __launch_bounds__(PT_THREADS, 1) void
k_tiles(pt_launch p)
{
	typedef Form<F> FM;
	typedef Lds<F> LD;
	typedef typename FM::G G;
	typedef typename FM::L L;
	typedef typename FM::AE AE;
	typedef typename FM::ACC ACC;
	typedef typename FM::ER ER;
	constexpr int KC = FM::KC;
	constexpr int NCB = (F == PT_FORM_FP32 && NSET == 1) ? 2 : 1; /* pt_tcw */
	constexpr int TCW = 16 * NCB;
	constexpr int NPR = 4;                     /* (u, w) pairs per thread */
	constexpr int SEPT = 32 * KC / PT_THREADS; /* S elements per thread */

	__shared__ __attribute__((aligned(16))) char smem[LD::BYTES];
	__shared__ double red[4];
	L *sS = (L *)smem;
	L *sQ = sS + FM::NP * LD::S_PLANE;
	AE *sX = (AE *)smem;

	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wa = wave & 1, wb = wave >> 1;
	const int ln = lane & 15, lg = lane >> 4;

	/* XCD-aware item order: workgroup n runs on XCD n % 8; each XCD gets a
	 * contiguous run of items, so neighbouring tiles of one ijk share L2 */
	const int64_t per = (p.n_chunk + 7) / 8;
	int64_t it = (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
	if (it >= p.n_chunk)
		return;
	it += p.item0;
	const int64_t ijk_id = it / p.n_tiles;
	const int tile = (int)(it - ijk_id * p.n_tiles);
	const int i = p.ijk[3 * ijk_id], j = p.ijk[3 * ijk_id + 1];
	const int k = p.ijk[3 * ijk_id + 2];
	const int A0 = p.tiles[3 * tile] * 32, B0 = p.tiles[3 * tile + 1] * 32;
	const int C0 = p.tiles[3 * tile + 2] * TCW;
	const int na = p.na, nc = p.nc;
	const int a_lo = A0 + 16 * wa, b_lo = B0 + 16 * wb;
	/* does this wave's block hold any element of the triangle? */
	const bool act = a_lo < na && b_lo < na && a_lo + 15 > b_lo &&
	    (KIND != PT_AAA || b_lo + 15 > C0);

	ACC acc[NSET][TCW];
#pragma unroll
	for (int s = 0; s < NSET; ++s)
#pragma unroll
		for (int x = 0; x < TCW; ++x)
			acc[s][x] = ACC{0, 0, 0, 0};

	/* steps = (term, kappa chunk) over all terms; phase 1 first */
	const int nterm = p.nterm1 + p.nterm2;
	G qr[NPR][KC], sr[SEPT];
	int t = 0, kc = 0;
	Step cur = make_step(p, p.terms[0], i, j, k, A0, B0, C0, TCW);
	int64_t soff = 0, qoff = 0;
	{
		const pt_term &T = p.terms[0];
		soff = T.s_off[0] * i + T.s_off[1] * j + T.s_off[2] * k;
		qoff = T.q_off[0] * i + T.q_off[1] * j + T.q_off[2] * k;
	}
	load_step<F, NPR, SEPT>(cur, soff, qoff, 0, tid, qr, sr);

	while (t < nterm) {
		/* phase boundary: [a][b]@c -> [a][c]@b, per register r */
		if (t == p.nterm1 && kc == 0) {
#pragma unroll
			for (int s = 0; s < NSET; ++s)
#pragma unroll
			for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				AE *X = sX + wave * (4 * 16 * 17);
				__syncthreads();
#pragma unroll
				for (int c = 0; c < 16; ++c)
					X[(lg * 16 + ln) * 17 + c] = acc[s][16 * cb + c][r];
				__syncthreads();
#pragma unroll
				for (int b = 0; b < 16; ++b)
					acc[s][16 * cb + b][r] = X[(lg * 16 + b) * 17 + ln];
			}
		}
		__syncthreads();
		store_step<F, NPR, SEPT>(cur, tid, sS, sQ, qr, sr);
		__syncthreads();
		/* prefetch the next step while this one computes */
		const int role = cur.role, set = cur.set;
		int tn = t, kn = kc + KC;
		if (kn >= cur.K) {
			tn = t + 1;
			kn = 0;
		}
		if (tn < nterm) {
			if (tn != t) {
				const pt_term &T = p.terms[tn];
				cur = make_step(p, T, i, j, k, A0, B0, C0, TCW);
				soff = T.s_off[0] * i + T.s_off[1] * j + T.s_off[2] * k;
				qoff = T.q_off[0] * i + T.q_off[1] * j + T.q_off[2] * k;
			}
			load_step<F, NPR, SEPT>(cur, soff, qoff, kn, tid, qr, sr);
		}
		if (act) {
			if (NSET == 1 || set == 0)
				mma_step<F, NCB>(acc[0], role, sS, sQ, wa, wb, ln, lg);
			else
				mma_step<F, NCB>(acc[NSET - 1], role, sS, sQ, wa, wb, ln,
				    lg);
		}
		t = tn;
		kc = kn;
	}

	/* epilogue: layout [a][c]@b, block x = 16 cb + b; lane holds
	 * c = C0 + 16 cb + ln.  Each register column r goes through this
	 * lane's own LDS slots so the element loop runs with a runtime index. */
	double esum = 0.0, ecmp = 0.0;
	AE *E = sX + wave * (NSET * TCW * 64) + lane;
#pragma unroll
	for (int r = 0; r < 4; ++r) {
		__syncthreads();
#pragma unroll
		for (int s = 0; s < NSET; ++s)
#pragma unroll
			for (int x = 0; x < TCW; ++x)
				E[(s * TCW + x) * 64] = acc[s][x][r];
		if (!act)
			continue;
		const int arow = FM::ROW_F64 ? (lg + 4 * r) : (4 * lg + r);
		const int a = a_lo + arow;
		if (a >= na)
			continue;
		for (int x = 0; x < TCW; ++x) {
			const int b = b_lo + (x & 15);
			const int c = C0 + 16 * (x >> 4) + ln;
			if (!(c < nc && a > b && (KIND != PT_AAA || b > c)))
				continue;
			ER z, dn;
			const ER *d1 = (const ER *)p.epi.d1;
			if (KIND == PT_AAA) {
				const int64_t o = p.oa, v = p.va;
				z = z_same<ER>((const ER *)p.epi.u1, (const ER *)p.epi.GA,
				    o, v, i, j, k, a, b, c) +
				    z_same<ER>((const ER *)p.epi.f1, (const ER *)p.epi.ZA,
				    o, v, i, j, k, a, b, c);
				dn = d1[i * v + a] + d1[j * v + b] + d1[k * v + c];
			} else {
				const ER *d2 = (const ER *)p.epi.d2;
				z = z_mixed<ER>(p.epi, p.oa, p.va, p.ob, p.vb, i, j, k, a, b,
				    c);
				dn = d1[i * p.va + a] + d1[j * p.va + b] + d2[k * p.vb + c];
			}
			ER term;
			if (NSET == 1) {
				ER w = (ER)E[x * 64];
				if (FM::SPLIT)
					w *= (ER)p.inv_scale[0];
				term = w * (w - z) / dn;
			} else {
				ER wl = (ER)E[x * 64], wr = (ER)E[(TCW + x) * 64];
				if (FM::SPLIT) {
					wl *= (ER)p.inv_scale[0];
					wr *= (ER)p.inv_scale[1];
				}
				term = (wl - z) * wr / dn;
			}
			kahan(esum, ecmp, (double)term);
		}
	}

	/* deterministic workgroup reduction */
	double e = esum - ecmp;
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1)
		e += __shfl_xor(e, off);
	if (lane == 0)
		red[wave] = e;
	__syncthreads();
	if (tid == 0)
		p.partial[it] = (red[0] + red[1]) + (red[2] + red[3]);
}

/* one workgroup, fixed-order compensated sum of x[0..n) */
__global__ __launch_bounds__(1024) void
k_reduce(const double *x, int64_t n, double *out)
{
	__shared__ double sh[1024];
	double s = 0.0, c = 0.0;
	for (int64_t t = threadIdx.x; t < n; t += 1024)
		kahan(s, c, x[t]);
	sh[threadIdx.x] = s - c;
	__syncthreads();
	for (int w = 512; w >= 1; w >>= 1) {
		if ((int)threadIdx.x < w)
			sh[threadIdx.x] += sh[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		out[0] = sh[0];
}

/* out[s] = sum of segment s (in order) */
__global__ void
k_segsum(const double *x, int64_t nseg, int64_t seglen, double *out)
{
	int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= nseg)
		return;
	double acc = 0.0, c = 0.0;
	for (int64_t t = 0; t < seglen; ++t)
		kahan(acc, c, x[s * seglen + t]);
	out[s] = acc - c;
}

/* hi + lo split of (x * scale) into bf16 or fp16 pairs */
__global__ void
k_split(const double *src, uint32_t *dst, int64_t n, int fp16, double scale)
{
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	    t += (int64_t)gridDim.x * blockDim.x) {
		double x = src[t] * scale;
		uint16_t hb, lb;
		if (fp16) {
			_Float16 h = (_Float16)x;
			_Float16 l = (_Float16)(x - (double)h);
			hb = __builtin_bit_cast(uint16_t, h);
			lb = __builtin_bit_cast(uint16_t, l);
		} else {
			__bf16 h = (__bf16)(float)x;
			__bf16 l = (__bf16)(float)(x - (double)(float)h);
			hb = __builtin_bit_cast(uint16_t, h);
			lb = __builtin_bit_cast(uint16_t, l);
		}
		dst[t] = (uint32_t)hb | ((uint32_t)lb << 16);
	}
}

/* max |x| as the bit pattern of a non-negative double (order-free) */
__global__ void
k_absmax(const double *src, int64_t n, unsigned long long *out)
{
	__shared__ double sh[256];
	double m = 0.0;
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	    t += (int64_t)gridDim.x * blockDim.x)
		m = fmax(m, fabs(src[t]));
	sh[threadIdx.x] = m;
	__syncthreads();
	for (int w = 128; w >= 1; w >>= 1) {
		if ((int)threadIdx.x < w)
			sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
		__syncthreads();
	}
	if (threadIdx.x == 0)
		atomicMax(out, (unsigned long long)__double_as_longlong(sh[0]));
}

__global__ void
k_synth(pt_synth_spec spec, int tensor, void *dst, int64_t n, int fp32)
{
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	    t += (int64_t)gridDim.x * blockDim.x) {
		double x = pt_synth_elem(&spec, tensor, t);
		if (fp32)
			((float *)dst)[t] = (float)x;
		else
			((double *)dst)[t] = x;
	}
}

template <int F>
int
launch_form(const pt_launch *p, hipStream_t s)
{
	dim3 grid((unsigned)(8 * ((p->n_chunk + 7) / 8))), block(PT_THREADS);
	if (p->kind == PT_AAA) {
		if (p->nset == 1)
			k_tiles<F, PT_AAA, 1><<<grid, block, 0, s>>>(*p);
		else
			k_tiles<F, PT_AAA, 2><<<grid, block, 0, s>>>(*p);
	} else {
		if (p->nset == 1)
			k_tiles<F, PT_AAB, 1><<<grid, block, 0, s>>>(*p);
		else
			k_tiles<F, PT_AAB, 2><<<grid, block, 0, s>>>(*p);
	}
	return (int)hipGetLastError();
}

} /* namespace */

extern "C" int
pt_launch_tiles(int form, const struct pt_launch *p, void *stream)
{
	hipStream_t s = (hipStream_t)stream;
	if (p->n_chunk <= 0)
		return 0;
	switch (form) {
	case PT_FORM_FP64: return launch_form<PT_FORM_FP64>(p, s);
	case PT_FORM_FP32: return launch_form<PT_FORM_FP32>(p, s);
	case PT_FORM_BF16X3: return launch_form<PT_FORM_BF16X3>(p, s);
	case PT_FORM_F16X3: return launch_form<PT_FORM_F16X3>(p, s);
	}
	return (int)hipErrorInvalidValue;
}

extern "C" int
pt_launch_reduce(const double *x, int64_t n, double *out, void *stream)
{
	k_reduce<<<1, 1024, 0, (hipStream_t)stream>>>(x, n, out);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_segsum(const double *x, int64_t nseg, int64_t seglen, double *out,
    void *stream)
{
	if (nseg <= 0)
		return 0;
	k_segsum<<<(unsigned)((nseg + 255) / 256), 256, 0, (hipStream_t)stream>>>(
	    x, nseg, seglen, out);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_split(const double *src, uint32_t *dst, int64_t n, int fp16,
    double scale, void *stream)
{
	if (n <= 0)
		return 0;
	int64_t nb = (n + 255) / 256;
	k_split<<<(unsigned)(nb < 65536 ? nb : 65536), 256, 0,
	    (hipStream_t)stream>>>(src, dst, n, fp16, scale);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_absmax(const double *src, int64_t n, double *out, void *stream)
{
	if (n <= 0)
		return 0;
	int64_t nb = (n + 255) / 256;
	k_absmax<<<(unsigned)(nb < 4096 ? nb : 4096), 256, 0,
	    (hipStream_t)stream>>>(src, n, (unsigned long long *)out);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_synth(const void *spec, int tensor, void *dst, int64_t n, int fp32,
    void *stream)
{
	if (n <= 0)
		return 0;
	int64_t nb = (n + 255) / 256;
	k_synth<<<(unsigned)(nb < 65536 ? nb : 65536), 256, 0,
	    (hipStream_t)stream>>>(*(const pt_synth_spec *)spec, tensor, dst, n,
	    fp32);
	return (int)hipGetLastError();
}
