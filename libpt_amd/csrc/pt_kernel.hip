/*
 * CDNA4 (gfx950) kernels of the perturbative-triples engine.
 *
 * k_tiles: one workgroup = one ijk item x one (A, B) tile of 32 a x 32 b,
 * looping over 16-wide c blocks.  Four waves each own a 16 x 16 x 16
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
	static constexpr int NP = 1, KC = 8, KCP = 9, KS = 4;
	static constexpr bool ROW_F64 = true, QC_HALF = false, SPLIT = false;
};

template <> struct Form<PT_FORM_FP32> {
	typedef float G;
	typedef float L;
	typedef float AE;
	typedef v4f ACC;
	typedef float ER;
	static constexpr int NP = 1, KC = 16, KCP = 17, KS = 4;
	static constexpr bool ROW_F64 = false, QC_HALF = false, SPLIT = false;
};

template <> struct Form<PT_FORM_BF16X3> {
	typedef uint32_t G;  /* hi in bits 0-15, lo in bits 16-31 */
	typedef uint16_t L;
	typedef float AE;
	typedef v4f ACC;
	typedef double ER;
	static constexpr int NP = 2, KC = 32, KCP = 40, KS = 32;
	static constexpr bool ROW_F64 = false, QC_HALF = true, SPLIT = true;
};

template <> struct Form<PT_FORM_F16X3> : Form<PT_FORM_BF16X3> {};

template <int F> struct Lds {
	typedef Form<F> FM;
	static constexpr int QWC = FM::QC_HALF ? 16 : 32;
	static constexpr int S_ELEMS = FM::NP * 32 * FM::KCP;
	static constexpr int Q_ELEMS = FM::NP * 32 * QWC * FM::KCP;
	static constexpr int STAGE_BYTES =
	    (S_ELEMS + Q_ELEMS) * (int)sizeof(typename FM::L);
	/* transpose scratch (4 waves x 4 x 16 x 17) and epilogue scratch
	 * (4 waves x 2 sets x 16 x 64) share the staging area */
	static constexpr int X_BYTES =
	    4 * 2 * 16 * 64 * (int)sizeof(typename FM::AE);
	static constexpr int BYTES =
	    STAGE_BYTES > X_BYTES ? STAGE_BYTES : X_BYTES;
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

/* ------------------------------------------------------------- staging */

template <class G>
__device__ __forceinline__ G
neg_if(G v, int neg)
{
	if constexpr (sizeof(G) == 4 && !__is_same(G, float))
		return neg ? (v ^ 0x80008000u) : v; /* both halves */
	else
		return neg ? -v : v;
}

template <int F>
__device__ __forceinline__ void
lds_put(typename Form<F>::L *base, int plane_elems, int idx,
    typename Form<F>::G v)
{
	if constexpr (Form<F>::SPLIT) {
		base[idx] = (uint16_t)(v & 0xffffu);
		base[plane_elems + idx] = (uint16_t)(v >> 16);
	} else {
		base[idx] = v;
	}
}

/* Role geometry of one term inside the workgroup tile. */
struct RoleGeom {
	int s0, sn, sext;   /* single index: origin, rows, extent */
	int u0, uext;       /* pair index u */
	int w0, wext, qw;   /* pair index w, LDS width */
	int urole, wrole;
};

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

/* MFMA over one staged kappa chunk into a fixed accumulator set */
template <int F>
__device__ __forceinline__ void
mma_chunk(typename Form<F>::ACC (&ac)[16], int role,
    const typename Form<F>::L *sS, const typename Form<F>::L *sQ, int wa,
    int wb, int ln, int lg)
{
	typedef Form<F> FM;
	typedef typename FM::L L;
	constexpr int KC = FM::KC, KCP = FM::KCP, KS = FM::KS;
	constexpr int QWC = Lds<F>::QWC;
	constexpr int SPE = 32 * KCP, QPE = 32 * QWC * KCP;
	(void)SPE; (void)QPE; (void)KS; (void)KC;
	if constexpr (!FM::SPLIT) {
#pragma unroll
		for (int ks = 0; ks < KC; ks += KS) {
			const int kap = ks + lg;
			if (role == PT_ROLE_A) {
				L av = sS[(16 * wa + ln) * KCP + kap];
#pragma unroll
				for (int c = 0; c < 16; ++c) {
					L bv = sQ[((16 * wb + ln) * 16 + c) * KCP + kap];
					ac[c] = mma(av, bv, ac[c]);
				}
			} else if (role == PT_ROLE_B) {
				L bv = sS[(16 * wb + ln) * KCP + kap];
#pragma unroll
				for (int c = 0; c < 16; ++c) {
					L av = sQ[((16 * wa + ln) * 16 + c) * KCP + kap];
					ac[c] = mma(av, bv, ac[c]);
				}
			} else {
				L bv = sS[ln * KCP + kap];
#pragma unroll
				for (int b = 0; b < 16; ++b) {
					L av = sQ[((16 * wa + ln) * QWC + 16 * wb + b) * KCP + kap];
					ac[b] = mma(av, bv, ac[b]);
				}
			}
		}
	} else {
		const int ko = 8 * lg;
		const int wq = QWC == 16 ? 0 : 16 * wb;
#define LD8(base, row) (*(const v8u *)((base) + (row) * KCP + ko))
		if (role == PT_ROLE_A) {
			const int rs = 16 * wa + ln;
			v8u ah = LD8(sS, rs), al = LD8(sS + SPE, rs);
#pragma unroll
			for (int c = 0; c < 16; ++c) {
				const int rq = (16 * wb + ln) * 16 + c;
				v8u bh = LD8(sQ, rq), bl = LD8(sQ + QPE, rq);
				ac[c] = mma8<F>(al, bh, ac[c]);
				ac[c] = mma8<F>(ah, bl, ac[c]);
				ac[c] = mma8<F>(ah, bh, ac[c]);
			}
		} else if (role == PT_ROLE_B) {
			const int rs = 16 * wb + ln;
			v8u bh = LD8(sS, rs), bl = LD8(sS + SPE, rs);
#pragma unroll
			for (int c = 0; c < 16; ++c) {
				const int rq = (16 * wa + ln) * 16 + c;
				v8u ah = LD8(sQ, rq), al = LD8(sQ + QPE, rq);
				ac[c] = mma8<F>(al, bh, ac[c]);
				ac[c] = mma8<F>(ah, bl, ac[c]);
				ac[c] = mma8<F>(ah, bh, ac[c]);
			}
		} else {
			const int rs = ln;
			v8u bh = LD8(sS, rs), bl = LD8(sS + SPE, rs);
#pragma unroll
			for (int b = 0; b < 16; ++b) {
				const int rq = (16 * wa + ln) * QWC + wq + b;
				v8u ah = LD8(sQ, rq), al = LD8(sQ + QPE, rq);
				ac[b] = mma8<F>(al, bh, ac[b]);
				ac[b] = mma8<F>(ah, bl, ac[b]);
				ac[b] = mma8<F>(ah, bh, ac[b]);
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
	typedef typename FM::G G;
	typedef typename FM::L L;
	typedef typename FM::AE AE;
	typedef typename FM::ACC ACC;
	typedef typename FM::ER ER;
	constexpr int NP = FM::NP, KC = FM::KC, KCP = FM::KCP, KS = FM::KS;
	constexpr int QWC = Lds<F>::QWC;
	constexpr int SPE = 32 * KCP;          /* S plane elements */
	constexpr int QPE = 32 * QWC * KCP;    /* Q plane elements */

	__shared__ __attribute__((aligned(16))) char smem[Lds<F>::BYTES];
	__shared__ double red[4];
	L *sS = (L *)smem;
	L *sQ = sS + NP * SPE;
	AE *sX = (AE *)smem;

	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wa = wave & 1, wb = wave >> 1;
	const int ln = lane & 15, lg = lane >> 4;

	/* XCD-aware item order: workgroup n runs on XCD n % 8; give each XCD
	 * a contiguous run of items so items sharing an ijk share its L2. */
	const int64_t per = (p.n_chunk + 7) / 8;
	int64_t it = (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
	if (it >= p.n_chunk)
		return;
	it += p.item0;
	const int64_t ijk_id = it / p.n_tiles;
	const int tile = (int)(it - ijk_id * p.n_tiles);
	const int i = p.ijk[3 * ijk_id], j = p.ijk[3 * ijk_id + 1];
	const int k = p.ijk[3 * ijk_id + 2];
	const int Ab = p.tiles[2 * tile], Bb = p.tiles[2 * tile + 1];
	const int A0 = Ab * 32, B0 = Bb * 32;
	const int na = p.na, nc = p.nc;
	const int cmax = KIND == PT_AAA ? min(2 * Bb + 1, p.nC - 1) : p.nC - 1;
	const int a_lo = A0 + 16 * wa, b_lo = B0 + 16 * wb;

	double esum = 0.0, ecmp = 0.0;

	for (int Cb = 0; Cb <= cmax; ++Cb) {
		const int C0 = Cb * 16;
		ACC acc[NSET][16];
#pragma unroll
		for (int s = 0; s < NSET; ++s)
#pragma unroll
			for (int x = 0; x < 16; ++x)
				acc[s][x] = ACC{0, 0, 0, 0};
		/* does this wave's 16^3 block hold any element of the triangle? */
		const bool act = a_lo < na && b_lo < na && a_lo + 15 > b_lo &&
		    (KIND != PT_AAA || b_lo + 15 > C0);

		for (int ph = 1; ph <= 2; ++ph) {
			const int t0 = ph == 1 ? 0 : p.nterm1;
			const int t1 = ph == 1 ? p.nterm1 : p.nterm1 + p.nterm2;
			if (ph == 2) {
				/* [a][b]@c -> [a][c]@b, per register r */
#pragma unroll
				for (int s = 0; s < NSET; ++s)
#pragma unroll
				for (int r = 0; r < 4; ++r) {
					AE *X = sX + wave * (4 * 16 * 17);
					__syncthreads();
#pragma unroll
					for (int c = 0; c < 16; ++c)
						X[(lg * 16 + ln) * 17 + c] = acc[s][c][r];
					__syncthreads();
#pragma unroll
					for (int b = 0; b < 16; ++b)
						acc[s][b][r] = X[(lg * 16 + b) * 17 + ln];
				}
			}
			for (int t = t0; t < t1; ++t) {
				const pt_term &T = p.terms[t];
				const int role = T.role, set = T.set, K = T.K;
				const int neg = T.sign < 0, packed = T.packed;
				const int ufast = T.ufast;
				const G *S = (const G *)p.buf[T.s_buf] +
				    T.s_off[0] * i + T.s_off[1] * j + T.s_off[2] * k;
				const G *Q = (const G *)p.buf[T.q_buf] +
				    T.q_off[0] * i + T.q_off[1] * j + T.q_off[2] * k;
				const int64_t s_st = T.s_st, q_k = T.q_k;
				RoleGeom g;
				if (role == PT_ROLE_A) {
					g = RoleGeom{A0, 32, na, B0, na, C0, nc, 16,
					    PT_ROLE_B, PT_ROLE_C};
				} else if (role == PT_ROLE_B) {
					g = RoleGeom{B0, 32, na, A0, na, C0, nc, 16,
					    PT_ROLE_A, PT_ROLE_C};
				} else {
					g = RoleGeom{C0, 16, nc, A0, na, B0, na, QWC,
					    PT_ROLE_A, PT_ROLE_B};
				}
				const int64_t q_u = T.q_st[g.urole], q_w = T.q_st[g.wrole];
				const int nhalf = (role == PT_ROLE_C && FM::QC_HALF) ? 2 : 1;

				for (int h = 0; h < nhalf; ++h) {
				const int w0 = g.w0 + 16 * h;
				for (int kc = 0; kc < K; kc += KC) {
					__syncthreads();
					/* S rows [s][kappa], sign folded in */
					for (int e = tid; e < 32 * KC; e += PT_THREADS) {
						const int s = e / KC, kk = e % KC;
						const int sg = g.s0 + s, kg = kc + kk;
						G v = G(0);
						if (s < g.sn && sg < g.sext && kg < K)
							v = S[(int64_t)sg * s_st + kg];
						lds_put<F>(sS, SPE, s * KCP + kk, neg_if(v, neg));
					}
					/* Q tile [u][w][kappa] */
					const int nq = 32 * g.qw * KC;
					const int qsh = g.qw == 32 ? 5 : 4;
					for (int e = tid; e < nq; e += PT_THREADS) {
						int u, w, kk;
						if (ufast) {
							u = e & 31;
							w = (e >> 5) & (g.qw - 1);
							kk = (e >> 5) >> qsh;
						} else {
							w = e & (g.qw - 1);
							u = (e >> qsh) & 31;
							kk = (e >> qsh) >> 5;
						}
						const int ug = g.u0 + u, wg = w0 + w, kg = kc + kk;
						G v = G(0);
						if (ug < g.uext && wg < g.wext && kg < K &&
						    (!packed || ug > wg)) {
							int64_t off = (int64_t)kg * q_k;
							if (packed)
								off += (int64_t)ug * (ug - 1) / 2 + wg;
							else
								off += (int64_t)ug * q_u +
								    (int64_t)wg * q_w;
							v = Q[off];
						}
						lds_put<F>(sQ, QPE, (u * g.qw + w) * KCP + kk, v);
					}
					__syncthreads();
					if (!act)
						continue;
					if (role == PT_ROLE_C && FM::QC_HALF && wb != h)
						continue;
					if (NSET == 1 || set == 0)
						mma_chunk<F>(acc[0], role, sS, sQ, wa, wb, ln, lg);
					else
						mma_chunk<F>(acc[NSET - 1], role, sS, sQ, wa, wb,
						    ln, lg);
				}
				}
			}
		}

		/* epilogue: layout [a][c]@b; lane holds c = C0 + ln.  Each
		 * register column r goes through this lane's own LDS slots so
		 * the element loop can run with a runtime index. */
		{
			const int c = C0 + ln;
			AE *E = sX + wave * (NSET * 16 * 64) + lane;
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				__syncthreads();
#pragma unroll
				for (int s = 0; s < NSET; ++s)
#pragma unroll
					for (int bb = 0; bb < 16; ++bb)
						E[(s * 16 + bb) * 64] = acc[s][bb][r];
				if (!act)
					continue;
				const int arow = FM::ROW_F64 ? (lg + 4 * r) : (4 * lg + r);
				const int a = a_lo + arow;
				if (a >= na || c >= nc)
					continue;
				for (int bb = 0; bb < 16; ++bb) {
					const int b = b_lo + bb;
					if (!(a > b && (KIND != PT_AAA || b > c)))
						continue;
					ER z, dn;
					const ER *d1 = (const ER *)p.epi.d1;
					if (KIND == PT_AAA) {
						const int64_t o = p.oa, v = p.va;
						z = z_same<ER>((const ER *)p.epi.u1,
						    (const ER *)p.epi.GA, o, v, i, j, k, a, b, c) +
						    z_same<ER>((const ER *)p.epi.f1,
						    (const ER *)p.epi.ZA, o, v, i, j, k, a, b, c);
						dn = d1[i * v + a] + d1[j * v + b] + d1[k * v + c];
					} else {
						const ER *d2 = (const ER *)p.epi.d2;
						z = z_mixed<ER>(p.epi, p.oa, p.va, p.ob, p.vb,
						    i, j, k, a, b, c);
						dn = d1[i * p.va + a] + d1[j * p.va + b] +
						    d2[k * p.vb + c];
					}
					ER term;
					if (NSET == 1) {
						ER w = (ER)E[bb * 64];
						if (FM::SPLIT)
							w *= (ER)p.inv_scale[0];
						term = w * (w - z) / dn;
					} else {
						ER wl = (ER)E[bb * 64];
						ER wr = (ER)E[(16 + bb) * 64];
						if (FM::SPLIT) {
							wl *= (ER)p.inv_scale[0];
							wr *= (ER)p.inv_scale[1];
						}
						term = (wl - z) * wr / dn;
					}
					kahan(esum, ecmp, (double)term);
				}
			}
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
