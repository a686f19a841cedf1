/*
 * Body of the CPU oracle, included twice by pt_oracle.c:
 *   R = double, FN(x) = x##_d   restates src/pt.c   (libpt FP64 engine)
 *   R = float,  FN(x) = x##_f   restates src/ptmp.c (libpt "_mp" engine:
 *                               float contractions, W, Z, Dn and per-element
 *                               term; double energy accumulator,
 *                               src/ptmp.c:192,285,428,543)
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links this.
 *
 * Every array follows the layouts of src/pt.h:43-57 (row-major, last index
 * fastest).  For one ijk triple the connected triples W[a][b][c] are built
 * from GEMM intermediates plus explicit per-element sums over the occupied
 * index, then combined with the disconnected term Z and the denominator Dn
 * in one elementwise energy pass.
 */

/* C[x*ldc + p] = sum_d A[x*lda + d] * B[d*ldb + p], x < m, p < n. */
static void
FN(mm)(size_t m, size_t n, size_t k, const R *A, size_t lda, const R *B,
    size_t ldb, R *C, size_t ldc)
{
	const size_t pbs = 512;
	size_t npb = (n + pbs - 1) / pbs;
	ptrdiff_t nwork = (ptrdiff_t)(npb * m), w;

	if (m == 0 || n == 0)
		return;
#pragma omp parallel for schedule(static)
	for (w = 0; w < nwork; w++) {
		size_t pb = (size_t)w / m, x = (size_t)w % m;
		size_t p0 = pb * pbs, p1 = p0 + pbs < n ? p0 + pbs : n, p, d;
		R *c = C + x * ldc;

		for (p = p0; p < p1; p++)
			c[p] = 0;
		for (d = 0; d < k; d++) {
			R s = A[x * lda + d];
			const R *b = B + d * ldb;
			for (p = p0; p < p1; p++)
				c[p] += s * b[p];
		}
	}
}

/* dst[(x*ny + y)*nl + l] = src[(l*nx + x)*ny + y]: moves the occupied index
 * of one T2 slab T[p][l][x][y] innermost so oovo sums are contiguous. */
static void
FN(lt)(size_t nl, size_t nx, size_t ny, const R *src, R *dst)
{
	ptrdiff_t x;

#pragma omp parallel for schedule(static)
	for (x = 0; x < (ptrdiff_t)nx; x++) {
		size_t y, l;
		for (y = 0; y < ny; y++)
			for (l = 0; l < nl; l++)
				dst[((size_t)x * ny + y) * nl + l] =
				    src[(l * nx + (size_t)x) * ny + y];
	}
}

static R
FN(dot)(const R *x, const R *y, size_t n)
{
	R s = 0;
	size_t l;

	for (l = 0; l < n; l++)
		s += x[l] * y[l];
	return s;
}

/* m[r][c] -> mt[c][r] */
static void
FN(tr)(size_t nr, size_t nc, const R *m, R *mt)
{
	ptrdiff_t c;

#pragma omp parallel for schedule(static)
	for (c = 0; c < (ptrdiff_t)nc; c++) {
		size_t r;
		for (r = 0; r < nr; r++)
			mt[(size_t)c * nr + r] = m[r * nc + (size_t)c];
	}
}

/*
 * Same-spin connected triples, src/pt.c:236-290 (and :588-696 for each of
 * the two (fT) builds).  Writes W[(a*v + b)*v + c] for a > b > c.
 *   Y_s[x][(y>z)] = sum_d T[p,q,x,d] V[r,d,(y,z)]
 *       for (p,q,r) = (i,j,k) +, (i,k,j) -, (k,j,i) -
 *   each contributes sign * (Y[c][ab] - Y[b][ac] + Y[a][bc])
 *   X_(p;q,r)(x,y,z) = sum_l T[p,l,x,y] O[q,r,z,l]
 *       for (p;q,r) = (i;j,k) +, (j;i,k) -, (k;j,i) -
 *   each contributes sign * (X(b,c,a) - X(a,c,b) - X(b,a,c))
 * scratch: 3*v*P(v) + 3*o*v*v elements.
 */
static void
FN(w_aaa)(size_t o, size_t v, const R *T, const R *O, const R *V, size_t i,
    size_t j, size_t k, R *W, R *scr)
{
	size_t P = v * (v - 1) / 2, vv = v * v;
	size_t pqr[3][3] = {{i, j, k}, {i, k, j}, {k, j, i}};
	size_t occ[3] = {i, j, k};
	/* (q,r) pairs paired with occ[n] as p */
	size_t qr[3][2] = {{j, k}, {i, k}, {j, i}};
	R *Y = scr, *Tt = scr + 3 * v * P;
	ptrdiff_t a;
	size_t s;

	for (s = 0; s < 3; s++)
		FN(mm)(v, P, v, T + (pqr[s][0] * o + pqr[s][1]) * vv, v,
		    V + pqr[s][2] * v * P, P, Y + s * v * P, P);
	for (s = 0; s < 3; s++)
		FN(lt)(o, v, v, T + occ[s] * o * vv, Tt + s * vv * o);

#pragma omp parallel for schedule(dynamic)
	for (a = 0; a < (ptrdiff_t)v; a++) {
		size_t b, c, n;
		for (b = 0; b < (size_t)a; b++)
		for (c = 0; c < b; c++) {
			size_t ab = (size_t)a * (a - 1) / 2 + b;
			size_t ac = (size_t)a * (a - 1) / 2 + c;
			size_t bc = b * (b - 1) / 2 + c;
			R w = 0;

			for (s = 0; s < 3; s++) {
				const R *y = Y + s * v * P;
				R t = y[c * P + ab] - y[b * P + ac] +
				    y[(size_t)a * P + bc];
				if (s == 0)
					w += t;
				else
					w -= t;
			}
			for (n = 0; n < 3; n++) {
				const R *tp = Tt + n * vv * o;
				const R *oq = O + (qr[n][0] * o + qr[n][1]) * v * o;
				R x1 = FN(dot)(tp + (b * v + c) * o,
				    oq + (size_t)a * o, o);
				R x2 = FN(dot)(tp + ((size_t)a * v + c) * o,
				    oq + b * o, o);
				R x3 = FN(dot)(tp + (b * v + (size_t)a) * o,
				    oq + c * o, o);
				R t = x1 - x2 - x3;
				if (n == 0)
					w += t;
				else
					w -= t;
			}
			W[((size_t)a * v + b) * v + c] = w;
		}
	}
}

/*
 * Mixed-spin connected triples, src/pt.c:361-432 (and :773-915).  Pair
 * spin sizes (oa,va), odd spin (ob,vb).  Writes W[(a*va + b)*vb + c], a > b.
 * Block roles: TA [oa,oa,va,va], TB [oa,ob,va,vb], TC [ob,oa,vb,va];
 * OA [oa,oa,va,oa], OB [oa,ob,va,ob], OC [ob,oa,vb,oa];
 * VA packed [oa,va,P(va)], VB [oa,vb,va,vb], VC [ob,va,vb,va].
 */
static void
FN(w_aab)(size_t oa, size_t va, size_t ob, size_t vb, const R *TA,
    const R *TB, const R *TC, const R *OA, const R *OB, const R *OC,
    const R *VA, const R *VB, const R *VC, size_t i, size_t j, size_t k,
    R *W, R *scr)
{
	size_t PA = va * (va - 1) / 2, nab = va * vb;
	R *H1 = scr;                 /* [a][c*va + b']  */
	R *H2 = H1 + va * nab;       /* [a][b'*vb + c]  */
	R *H3 = H2 + va * nab;
	R *H4 = H3 + va * nab;       /* [c][p] then transposed */
	R *G4a = H4 + vb * PA;       /* [p][c] */
	R *G4b = G4a + vb * PA;
	R *TAi = G4b + vb * PA;      /* [x][y][l], l over oa */
	R *TAj = TAi + va * va * oa;
	R *TBi = TAj + va * va * oa; /* [x][y][l], l over ob */
	R *TBj = TBi + nab * ob;
	R *TCk = TBj + nab * ob;     /* [x=c][y=a][l], l over oa */
	const R *OCkj = OC + (k * oa + j) * vb * oa;
	const R *OCki = OC + (k * oa + i) * vb * oa;
	const R *OBjk = OB + (j * ob + k) * va * ob;
	const R *OBik = OB + (i * ob + k) * va * ob;
	const R *OAji = OA + (j * oa + i) * va * oa;
	ptrdiff_t a;

	FN(mm)(va, nab, va, TA + (i * oa + j) * va * va, va,
	    VC + k * va * vb * va, nab, H1, nab);
	FN(mm)(va, nab, vb, TB + (i * ob + k) * nab, vb,
	    VB + j * vb * nab, nab, H2, nab);
	FN(mm)(va, nab, vb, TB + (j * ob + k) * nab, vb,
	    VB + i * vb * nab, nab, H3, nab);
	FN(mm)(vb, PA, va, TC + (k * oa + j) * vb * va, va,
	    VA + i * va * PA, PA, H4, PA);
	FN(tr)(vb, PA, H4, G4a);
	FN(mm)(vb, PA, va, TC + (k * oa + i) * vb * va, va,
	    VA + j * va * PA, PA, H4, PA);
	FN(tr)(vb, PA, H4, G4b);
	FN(lt)(oa, va, va, TA + i * oa * va * va, TAi);
	FN(lt)(oa, va, va, TA + j * oa * va * va, TAj);
	FN(lt)(ob, va, vb, TB + i * ob * nab, TBi);
	FN(lt)(ob, va, vb, TB + j * ob * nab, TBj);
	FN(lt)(oa, vb, va, TC + k * oa * vb * va, TCk);

#pragma omp parallel for schedule(dynamic)
	for (a = 0; a < (ptrdiff_t)va; a++) {
		size_t b, c;
		for (b = 0; b < (size_t)a; b++) {
			size_t ab = (size_t)a * (a - 1) / 2 + b;
			for (c = 0; c < vb; c++) {
				R w;

				w = -H1[(size_t)a * nab + c * va + b] +
				    H1[b * nab + c * va + (size_t)a];
				w += -H2[(size_t)a * nab + b * vb + c] +
				    H2[b * nab + (size_t)a * vb + c];
				w += H3[(size_t)a * nab + b * vb + c] -
				    H3[b * nab + (size_t)a * vb + c];
				w += -G4a[ab * vb + c] + G4b[ab * vb + c];
				w += -FN(dot)(TAi + ((size_t)a * va + b) * oa,
				    OCkj + c * oa, oa) +
				    FN(dot)(TAj + ((size_t)a * va + b) * oa,
				    OCki + c * oa, oa);
				w += -FN(dot)(TBi + ((size_t)a * vb + c) * ob,
				    OBjk + b * ob, ob) +
				    FN(dot)(TBi + (b * vb + c) * ob,
				    OBjk + (size_t)a * ob, ob);
				w += -FN(dot)(TBj + (b * vb + c) * ob,
				    OBik + (size_t)a * ob, ob) +
				    FN(dot)(TBj + ((size_t)a * vb + c) * ob,
				    OBik + b * ob, ob);
				w += -FN(dot)(TCk + (c * va + (size_t)a) * oa,
				    OAji + b * oa, oa) +
				    FN(dot)(TCk + (c * va + b) * oa,
				    OAji + (size_t)a * oa, oa);
				W[((size_t)a * va + b) * vb + c] = w;
			}
		}
	}
}

static size_t
FN(scr_aab)(size_t oa, size_t va, size_t ob, size_t vb)
{
	size_t PA = va * (va - 1) / 2, nab = va * vb;
	return 3 * va * nab + 3 * vb * PA + 2 * va * va * oa + 2 * nab * ob +
	    vb * va * oa;
}

/* Same-spin disconnected term, src/pt.c:162-175. */
static R
FN(zs)(size_t o, size_t v, const R *u, const R *g, size_t i, size_t j,
    size_t k, size_t a, size_t b, size_t c)
{
	size_t vv = v * v;
	const R *gjk = g + (j * o + k) * vv, *gik = g + (i * o + k) * vv;
	const R *gji = g + (j * o + i) * vv;

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

/* One mixed-spin disconnected product, src/pt.c:177-185. */
static R
FN(zb)(size_t v1, size_t o2, size_t v2a, size_t v2b, size_t i, size_t j,
    size_t k, size_t a, size_t b, size_t c, const R *u, const R *g,
    const R *f, const R *t)
{
	size_t x = ((j * o2 + k) * v2a + b) * v2b + c;
	return u[i * v1 + a] * g[x] + f[i * v1 + a] * t[x];
}

/* Compensated accumulation (Neumaier). */
struct FN(acc) { double s, c; };

static void
FN(add)(struct FN(acc) *acc, double x)
{
	double t = acc->s + x;
	if (fabs(acc->s) >= fabs(x))
		acc->c += (acc->s - t) + x;
	else
		acc->c += (x - t) + acc->s;
	acc->s = t;
}

/*
 * Same-spin (T) energy over a list of ijk items, src/pt.c:187-302.  items
 * holds 3*nitems occupied indices (i<j<k); NULL means every i<j<k in
 * lexicographic order.  e_items (nullable) receives each item's energy.
 */
double
FN(oracle_pt_aaa)(size_t o, size_t v, const R *d, const R *f, const R *t1,
    const R *T, const R *O, const R *G, const R *V, size_t nitems,
    const int32_t *items, double *e_items)
{
	size_t n, P = v * (v - 1) / 2;
	int32_t *all = NULL;
	R *W, *scr;
	double *erow;
	struct FN(acc) tot = {0, 0};

	if (o == 0 || v == 0)
		return 0.0;
	if (items == NULL) {
		nitems = o * (o - 1) * (o - 2) / 6;
		all = ora_ijk_list(o, o, 1, nitems);
		items = all;
	}
	if (nitems == 0) {
		free(all);
		return 0.0;
	}
	W = ora_alloc(v * v * v * sizeof(R));
	scr = ora_alloc((3 * v * P + 3 * o * v * v) * sizeof(R));
	erow = ora_alloc(v * sizeof(double));
	for (n = 0; n < nitems; n++) {
		size_t i = (size_t)items[3 * n], j = (size_t)items[3 * n + 1];
		size_t k = (size_t)items[3 * n + 2];
		struct FN(acc) e = {0, 0};
		ptrdiff_t a;

		FN(w_aaa)(o, v, T, O, V, i, j, k, W, scr);
#pragma omp parallel for schedule(dynamic)
		for (a = 0; a < (ptrdiff_t)v; a++) {
			struct FN(acc) ea = {0, 0};
			size_t b, c;
			for (b = 0; b < (size_t)a; b++)
			for (c = 0; c < b; c++) {
				R dn = d[i * v + a] + d[j * v + b] + d[k * v + c];
				R z = FN(zs)(o, v, t1, G, i, j, k, a, b, c) +
				    FN(zs)(o, v, f, T, i, j, k, a, b, c);
				R w = W[((size_t)a * v + b) * v + c];
				R term = w * (w - z) / dn;
				FN(add)(&ea, (double)term);
			}
			erow[a] = ea.s + ea.c;
		}
		for (a = 0; a < (ptrdiff_t)v; a++)
			FN(add)(&e, erow[a]);
		if (e_items)
			e_items[n] = e.s + e.c;
		FN(add)(&tot, e.s + e.c);
	}
	free(W);
	free(scr);
	free(erow);
	free(all);
	return tot.s + tot.c;
}

/*
 * Mixed-spin (T) energy, src/pt.c:304-452.  items: i<j<oa, k<ob; NULL means
 * all of them with k fastest.
 */
double
FN(oracle_pt_aab)(size_t oa, size_t va, size_t ob, size_t vb,
    const R *d1, const R *d2, const R *f1, const R *f2, const R *u1,
    const R *u2, const R *TA, const R *TB, const R *TC, const R *OA,
    const R *OB, const R *OC, const R *GA, const R *GB, const R *VA,
    const R *VB, const R *VC, size_t nitems, const int32_t *items,
    double *e_items)
{
	size_t n;
	int32_t *all = NULL;
	R *W, *scr;
	double *erow;
	struct FN(acc) tot = {0, 0};

	if (oa == 0 || va == 0 || ob == 0 || vb == 0)
		return 0.0;
	if (items == NULL) {
		nitems = oa * (oa - 1) / 2 * ob;
		all = ora_ijk_list(oa, ob, 0, nitems);
		items = all;
	}
	if (nitems == 0) {
		free(all);
		return 0.0;
	}
	W = ora_alloc(va * va * vb * sizeof(R));
	scr = ora_alloc(FN(scr_aab)(oa, va, ob, vb) * sizeof(R));
	erow = ora_alloc(va * sizeof(double));
	for (n = 0; n < nitems; n++) {
		size_t i = (size_t)items[3 * n], j = (size_t)items[3 * n + 1];
		size_t k = (size_t)items[3 * n + 2];
		struct FN(acc) e = {0, 0};
		ptrdiff_t a;

		FN(w_aab)(oa, va, ob, vb, TA, TB, TC, OA, OB, OC, VA, VB, VC,
		    i, j, k, W, scr);
#pragma omp parallel for schedule(dynamic)
		for (a = 0; a < (ptrdiff_t)va; a++) {
			struct FN(acc) ea = {0, 0};
			size_t b, c;
			for (b = 0; b < (size_t)a; b++)
			for (c = 0; c < vb; c++) {
				R z = +FN(zb)(va, ob, va, vb, i, j, k, a, b, c,
				        u1, GB, f1, TB)
				      -FN(zb)(va, ob, va, vb, i, j, k, b, a, c,
				        u1, GB, f1, TB)
				      -FN(zb)(va, ob, va, vb, j, i, k, a, b, c,
				        u1, GB, f1, TB)
				      +FN(zb)(va, ob, va, vb, j, i, k, b, a, c,
				        u1, GB, f1, TB)
				      +FN(zb)(vb, oa, va, va, k, j, i, c, b, a,
				        u2, GA, f2, TA);
				R dn = d1[i * va + a] + d1[j * va + b] +
				    d2[k * vb + c];
				R w = W[((size_t)a * va + b) * vb + c];
				R term = w * (w - z) / dn;
				FN(add)(&ea, (double)term);
			}
			erow[a] = ea.s + ea.c;
		}
		for (a = 0; a < (ptrdiff_t)va; a++)
			FN(add)(&e, erow[a]);
		if (e_items)
			e_items[n] = e.s + e.c;
		FN(add)(&tot, e.s + e.c);
	}
	free(W);
	free(scr);
	free(erow);
	free(all);
	return tot.s + tot.c;
}

/*
 * Same-spin (fT) energy, src/pt.c:537-707: W_L from (l2, i7, i6), W_R from
 * (t2, i3, i2), Z_L = Zs(l1, i_oovv) + Zs(f2, l2), E += (W_L - Z_L) W_R / Dn.
 */
double
FN(oracle_ft_aaa)(size_t o, size_t v, const R *d, const R *f2, const R *l1,
    const R *T, const R *L, const R *G, const R *I2, const R *I3,
    const R *I6, const R *I7, size_t nitems, const int32_t *items,
    double *e_items)
{
	size_t n, P = v * (v - 1) / 2;
	int32_t *all = NULL;
	R *WL, *WR, *scr;
	double *erow;
	struct FN(acc) tot = {0, 0};

	if (o == 0 || v == 0)
		return 0.0;
	if (items == NULL) {
		nitems = o * (o - 1) * (o - 2) / 6;
		all = ora_ijk_list(o, o, 1, nitems);
		items = all;
	}
	if (nitems == 0) {
		free(all);
		return 0.0;
	}
	WL = ora_alloc(v * v * v * sizeof(R));
	WR = ora_alloc(v * v * v * sizeof(R));
	scr = ora_alloc((3 * v * P + 3 * o * v * v) * sizeof(R));
	erow = ora_alloc(v * sizeof(double));
	for (n = 0; n < nitems; n++) {
		size_t i = (size_t)items[3 * n], j = (size_t)items[3 * n + 1];
		size_t k = (size_t)items[3 * n + 2];
		struct FN(acc) e = {0, 0};
		ptrdiff_t a;

		FN(w_aaa)(o, v, L, I6, I7, i, j, k, WL, scr);
		FN(w_aaa)(o, v, T, I2, I3, i, j, k, WR, scr);
#pragma omp parallel for schedule(dynamic)
		for (a = 0; a < (ptrdiff_t)v; a++) {
			struct FN(acc) ea = {0, 0};
			size_t b, c;
			for (b = 0; b < (size_t)a; b++)
			for (c = 0; c < b; c++) {
				size_t x = ((size_t)a * v + b) * v + c;
				R dn = d[i * v + a] + d[j * v + b] + d[k * v + c];
				R z = FN(zs)(o, v, l1, G, i, j, k, a, b, c) +
				    FN(zs)(o, v, f2, L, i, j, k, a, b, c);
				R term = (WL[x] - z) * WR[x] / dn;
				FN(add)(&ea, (double)term);
			}
			erow[a] = ea.s + ea.c;
		}
		for (a = 0; a < (ptrdiff_t)v; a++)
			FN(add)(&e, erow[a]);
		if (e_items)
			e_items[n] = e.s + e.c;
		FN(add)(&tot, e.s + e.c);
	}
	free(WL);
	free(WR);
	free(scr);
	free(erow);
	free(all);
	return tot.s + tot.c;
}

/* Mixed-spin (fT) energy, src/pt.c:709-934. */
double
FN(oracle_ft_aab)(size_t oa, size_t va, size_t ob, size_t vb,
    const R *d1, const R *d2, const R *f1, const R *f2, const R *u1,
    const R *u2, const R *TA, const R *TB, const R *TC, const R *LA,
    const R *LB, const R *LC, const R *GA, const R *GB, const R *I2A,
    const R *I2B, const R *I2C, const R *I3A, const R *I3B, const R *I3C,
    const R *I6A, const R *I6B, const R *I6C, const R *I7A, const R *I7B,
    const R *I7C, size_t nitems, const int32_t *items, double *e_items)
{
	size_t n;
	int32_t *all = NULL;
	R *WL, *WR, *scr;
	double *erow;
	struct FN(acc) tot = {0, 0};

	if (oa == 0 || va == 0 || ob == 0 || vb == 0)
		return 0.0;
	if (items == NULL) {
		nitems = oa * (oa - 1) / 2 * ob;
		all = ora_ijk_list(oa, ob, 0, nitems);
		items = all;
	}
	if (nitems == 0) {
		free(all);
		return 0.0;
	}
	WL = ora_alloc(va * va * vb * sizeof(R));
	WR = ora_alloc(va * va * vb * sizeof(R));
	scr = ora_alloc(FN(scr_aab)(oa, va, ob, vb) * sizeof(R));
	erow = ora_alloc(va * sizeof(double));
	for (n = 0; n < nitems; n++) {
		size_t i = (size_t)items[3 * n], j = (size_t)items[3 * n + 1];
		size_t k = (size_t)items[3 * n + 2];
		struct FN(acc) e = {0, 0};
		ptrdiff_t a;

		FN(w_aab)(oa, va, ob, vb, LA, LB, LC, I6A, I6B, I6C, I7A, I7B,
		    I7C, i, j, k, WL, scr);
		FN(w_aab)(oa, va, ob, vb, TA, TB, TC, I2A, I2B, I2C, I3A, I3B,
		    I3C, i, j, k, WR, scr);
#pragma omp parallel for schedule(dynamic)
		for (a = 0; a < (ptrdiff_t)va; a++) {
			struct FN(acc) ea = {0, 0};
			size_t b, c;
			for (b = 0; b < (size_t)a; b++)
			for (c = 0; c < vb; c++) {
				size_t x = ((size_t)a * va + b) * vb + c;
				R z = +FN(zb)(va, ob, va, vb, i, j, k, a, b, c,
				        u1, GB, f1, LB)
				      -FN(zb)(va, ob, va, vb, i, j, k, b, a, c,
				        u1, GB, f1, LB)
				      -FN(zb)(va, ob, va, vb, j, i, k, a, b, c,
				        u1, GB, f1, LB)
				      +FN(zb)(va, ob, va, vb, j, i, k, b, a, c,
				        u1, GB, f1, LB)
				      +FN(zb)(vb, oa, va, va, k, j, i, c, b, a,
				        u2, GA, f2, LA);
				R dn = d1[i * va + a] + d1[j * va + b] +
				    d2[k * vb + c];
				R term = (WL[x] - z) * WR[x] / dn;
				FN(add)(&ea, (double)term);
			}
			erow[a] = ea.s + ea.c;
		}
		for (a = 0; a < (ptrdiff_t)va; a++)
			FN(add)(&e, erow[a]);
		if (e_items)
			e_items[n] = e.s + e.c;
		FN(add)(&tot, e.s + e.c);
	}
	free(WL);
	free(WR);
	free(scr);
	free(erow);
	free(all);
	return tot.s + tot.c;
}

/* Block offsets and spin-case sums, src/pt.c:454-535 and :936-1050. */
double
FN(oracle_rpt)(size_t o, size_t v, const R *d, const R *f, const R *t1,
    const R *t2, const R *oovo, const R *oovv, const R *ovvv)
{
	const R *tA = t2, *tB = t2 + o * o * v * v;
	const R *oA = oovo, *oB = oovo + o * o * o * v;
	const R *gA = oovv, *gB = oovv + o * o * v * v;
	const R *vA = ovvv, *vB = ovvv + o * v * v * (v - 1) / 2;
	double e = 0;

	e += FN(oracle_pt_aaa)(o, v, d, f, t1, tA, oA, gA, vA, 0, NULL, NULL);
	e += FN(oracle_pt_aab)(o, v, o, v, d, d, f, f, t1, t1, tA, tB, tB, oA,
	    oB, oB, gA, gB, vA, vB, vB, 0, NULL, NULL);
	return 2.0 * e;
}

double
FN(oracle_upt)(size_t oa, size_t va, size_t ob, size_t vb, const R *d,
    const R *f, const R *t1, const R *t2, const R *oovo, const R *oovv,
    const R *ovvv)
{
	struct ora_blocks s = ora_split(oa, va, ob, vb);
	const R *db = d + s.ov, *fb = f + s.ov, *ub = t1 + s.ov;
	const R *tA = t2, *tB = t2 + s.oovv[0], *tBB = tB + s.oovv[1];
	const R *tC = tBB + s.oovv[2];
	const R *oA = oovo, *oB = oovo + s.oovo[0], *oBB = oB + s.oovo[1];
	const R *oC = oBB + s.oovo[2];
	const R *gA = oovv, *gB = oovv + s.oovv[0], *gBB = gB + s.oovv[1];
	const R *gC = gBB + s.oovv[2];
	const R *vA = ovvv, *vB = ovvv + s.ovvv[0], *vBB = vB + s.ovvv[1];
	const R *vC = vBB + s.ovvv[2];
	double e = 0;

	e += FN(oracle_pt_aaa)(oa, va, d, f, t1, tA, oA, gA, vA, 0, NULL,
	    NULL);
	e += FN(oracle_pt_aaa)(ob, vb, db, fb, ub, tBB, oBB, gBB, vBB, 0, NULL,
	    NULL);
	e += FN(oracle_pt_aab)(oa, va, ob, vb, d, db, f, fb, t1, ub, tA, tB,
	    tC, oA, oB, oC, gA, gB, vA, vB, vC, 0, NULL, NULL);
	e += FN(oracle_pt_aab)(ob, vb, oa, va, db, d, fb, f, ub, t1, tBB, tC,
	    tB, oBB, oC, oB, gBB, gC, vBB, vC, vB, 0, NULL, NULL);
	return e;
}

double
FN(oracle_rft)(size_t o, size_t v, const R *d, const R *f2, const R *l1,
    const R *t2, const R *l2, const R *oovv, const R *i2, const R *i3,
    const R *i6, const R *i7)
{
	size_t n2 = o * o * v * v, n3 = o * o * o * v;
	size_t np = o * v * v * (v - 1) / 2;
	double e = 0;

	e += FN(oracle_ft_aaa)(o, v, d, f2, l1, t2, l2, oovv, i2, i3, i6, i7,
	    0, NULL, NULL);
	e += FN(oracle_ft_aab)(o, v, o, v, d, d, f2, f2, l1, l1, t2, t2 + n2,
	    t2 + n2, l2, l2 + n2, l2 + n2, oovv, oovv + n2, i2, i2 + n3,
	    i2 + n3, i3, i3 + np, i3 + np, i6, i6 + n3, i6 + n3, i7, i7 + np,
	    i7 + np, 0, NULL, NULL);
	return 2.0 * e;
}

double
FN(oracle_uft)(size_t oa, size_t va, size_t ob, size_t vb, const R *d,
    const R *f2, const R *l1, const R *t2, const R *l2, const R *oovv,
    const R *i2, const R *i3, const R *i6, const R *i7)
{
	struct ora_blocks s = ora_split(oa, va, ob, vb);
	const R *db = d + s.ov, *fb = f2 + s.ov, *ub = l1 + s.ov;
	const R *t[4], *l[4], *g[4], *x2[4], *x3[4], *x6[4], *x7[4];
	double e = 0;
	int q;

	t[0] = t2; l[0] = l2; g[0] = oovv;
	x2[0] = i2; x6[0] = i6; x3[0] = i3; x7[0] = i7;
	for (q = 1; q < 4; q++) {
		t[q] = t[q - 1] + s.oovv[q - 1];
		l[q] = l[q - 1] + s.oovv[q - 1];
		g[q] = g[q - 1] + s.oovv[q - 1];
		x2[q] = x2[q - 1] + s.oovo[q - 1];
		x6[q] = x6[q - 1] + s.oovo[q - 1];
		x3[q] = x3[q - 1] + s.ovvv[q - 1];
		x7[q] = x7[q - 1] + s.ovvv[q - 1];
	}
	/* block order aaaa, abab, bbbb, baba = 0,1,2,3 */
	e += FN(oracle_ft_aaa)(oa, va, d, f2, l1, t[0], l[0], g[0], x2[0],
	    x3[0], x6[0], x7[0], 0, NULL, NULL);
	e += FN(oracle_ft_aaa)(ob, vb, db, fb, ub, t[2], l[2], g[2], x2[2],
	    x3[2], x6[2], x7[2], 0, NULL, NULL);
	e += FN(oracle_ft_aab)(oa, va, ob, vb, d, db, f2, fb, l1, ub, t[0],
	    t[1], t[3], l[0], l[1], l[3], g[0], g[1], x2[0], x2[1], x2[3],
	    x3[0], x3[1], x3[3], x6[0], x6[1], x6[3], x7[0], x7[1], x7[3], 0,
	    NULL, NULL);
	e += FN(oracle_ft_aab)(ob, vb, oa, va, db, d, fb, f2, ub, l1, t[2],
	    t[3], t[1], l[2], l[3], l[1], g[2], g[3], x2[2], x2[3], x2[1],
	    x3[2], x3[3], x3[1], x6[2], x6[3], x6[1], x7[2], x7[3], x7[1], 0,
	    NULL, NULL);
	return e;
}
