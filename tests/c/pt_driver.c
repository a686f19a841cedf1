/*
 * C caller of the pt.h drop-in: the way the reference's own test driver
 * (test.cpp:752-957) and Q-Chem use libpt, written against include/pt.h
 * and linked to libpt_amd.so with no other dependency.
 *
 *   pt_driver <rpt|upt|rft|uft> <mp 0|1> <oa> <va> [<ob> <vb>] <array files>
 *
 * Each array file holds the raw float64 values of one pt.h argument, in
 * pt.h order (7 arrays for (T), 10 for (fT)); with mp = 1 they are
 * converted to float and the _mp function is called.  The scratch hooks are
 * replaced by counting wrappers around malloc/free.  Prints
 *   E <energy, %.17g>
 *   hooks <malloc calls> <free calls>
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pt.h"

static long n_malloc, n_free;

static void *
count_malloc(size_t n)
{
	++n_malloc;
	return malloc(n);
}

static void
count_free(void *p)
{
	++n_free;
	free(p);
}

/* whole file as float64 (mp = 0) or float (mp = 1); NULL on error */
static void *
read_array(const char *path, int mp, size_t *count)
{
	FILE *f = fopen(path, "rb");
	if (!f)
		return NULL;
	fseek(f, 0, SEEK_END);
	long bytes = ftell(f);
	fseek(f, 0, SEEK_SET);
	size_t n = (size_t)bytes / sizeof(double);
	double *d = malloc(n ? n * sizeof(double) : 8);
	if (!d || fread(d, sizeof(double), n, f) != n) {
		fclose(f);
		free(d);
		return NULL;
	}
	fclose(f);
	*count = n;
	if (!mp)
		return d;
	float *s = malloc(n ? n * sizeof(float) : 4);
	if (!s) {
		free(d);
		return NULL;
	}
	for (size_t q = 0; q < n; ++q)
		s[q] = (float)d[q];
	free(d);
	return s;
}

int
main(int argc, char **argv)
{
	if (argc < 5) {
		fprintf(stderr, "usage: %s kind mp oa va [ob vb] files...\n",
		    argv[0]);
		return 2;
	}
	const char *kind = argv[1];
	const int mp = atoi(argv[2]);
	const int unres = kind[0] == 'u';
	const int ft = strcmp(kind + 1, "ft") == 0;
	const int nsz = unres ? 4 : 2, narr = ft ? 10 : 7;
	if (argc != 3 + nsz + narr) {
		fprintf(stderr, "expected %d sizes and %d arrays\n", nsz, narr);
		return 2;
	}
	size_t sz[4];
	for (int q = 0; q < nsz; ++q)
		sz[q] = (size_t)strtoull(argv[3 + q], NULL, 10);
	void *a[10];
	for (int q = 0; q < narr; ++q) {
		size_t n;
		a[q] = read_array(argv[3 + nsz + q], mp, &n);
		if (!a[q]) {
			fprintf(stderr, "cannot read %s\n", argv[3 + nsz + q]);
			return 2;
		}
	}
	libpt_set_malloc(count_malloc);
	libpt_set_free(count_free);
	double e = 0;
#define D(q) ((const double *)a[q])
#define S(q) ((const float *)a[q])
	if (!unres && !ft)
		e = mp ? libpt_rpt_mp(sz[0], sz[1], S(0), S(1), S(2), S(3), S(4),
		    S(5), S(6)) : libpt_rpt(sz[0], sz[1], D(0), D(1), D(2), D(3),
		    D(4), D(5), D(6));
	else if (unres && !ft)
		e = mp ? libpt_upt_mp(sz[0], sz[1], sz[2], sz[3], S(0), S(1),
		    S(2), S(3), S(4), S(5), S(6)) : libpt_upt(sz[0], sz[1], sz[2],
		    sz[3], D(0), D(1), D(2), D(3), D(4), D(5), D(6));
	else if (!unres)
		e = mp ? libpt_rft_mp(sz[0], sz[1], S(0), S(1), S(2), S(3), S(4),
		    S(5), S(6), S(7), S(8), S(9)) : libpt_rft(sz[0], sz[1], D(0),
		    D(1), D(2), D(3), D(4), D(5), D(6), D(7), D(8), D(9));
	else
		e = mp ? libpt_uft_mp(sz[0], sz[1], sz[2], sz[3], S(0), S(1),
		    S(2), S(3), S(4), S(5), S(6), S(7), S(8), S(9)) :
		    libpt_uft(sz[0], sz[1], sz[2], sz[3], D(0), D(1), D(2), D(3),
		    D(4), D(5), D(6), D(7), D(8), D(9));
#undef D
#undef S
	libpt_set_malloc(NULL);
	libpt_set_free(NULL);
	printf("E %.17g\n", e);
	printf("hooks %ld %ld\n", n_malloc, n_free);
	for (int q = 0; q < narr; ++q)
		free(a[q]);
	return 0;
}
