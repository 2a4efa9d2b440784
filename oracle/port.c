/*
 * C restatement of the reference CPU triples path (closed-shell (T)).
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/ as the checker, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the product
 * path (libcc_amd/, include/) never links or calls it.
 *
 * Algorithm, step for step, as the reference does it (file:line into the
 * reference tree):
 *   driver          src/triples-ijk/triples.F90:29-320 (tuple order, zeroing, weights, ets)
 *   task decode     src/triples-ijk/common_functions.f:62-86 (DDCC_T_TASK)
 *   VE fetch        src/triples-ijk/ddcc_aux.f:18-45 + trant3_1 => VE_x(d,p,q) = VVVO(d,q,p,x)
 *   W, i>j>k        src/triples-ijk/ijk_tuple.f90:56-84 (12 DGEMM, trant3_1/4/1/1, zerot3)
 *   E[T] + v1       src/triples-ijk/t1wt3.f90:123-192
 *   W, (i,j,j)      src/triples-ijk/ijj_tuple.f90:31-56 then ddcc_t_ijj_iij_big.f:330-342
 *   W, (i,i,j)      src/triples-ijk/iij_tuple.f90:95-123 then ddcc_t_ijj_iij_big.f:370-382
 *   helpers         src/triples-ijk/ddcc_aux.f:48-228, t3squa_acc.f:144-182
 *   permutations    src/common/common_cc.f90:62-333
 * All arrays are Fortran column-major; index macros below are 0-based.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

/* ------------------------------------------------------------------------ */
/* column-major GEMM, textbook definition: C = alpha*op(A)*op(B) + beta*C     */
/* ------------------------------------------------------------------------ */
static void gemm(char ta, char tb, i64 m, i64 n, i64 k, double alpha,
                 const double *A, i64 lda, const double *B, i64 ldb,
                 double beta, double *C, i64 ldc)
{
    double *at = NULL;
    const double *ap = A;
    i64 ldap = lda;
    if (ta == 't') {               /* materialise op(A) = A^T as m x k, ld m */
        at = (double *)malloc(sizeof(double) * (size_t)(m * k));
        for (i64 l = 0; l < k; ++l)
            for (i64 i = 0; i < m; ++i)
                at[i + m * l] = A[l + lda * i];
        ap = at;
        ldap = m;
    }
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < n; ++j) {
        double *cj = C + ldc * j;
        if (beta == 0.0) {
            for (i64 i = 0; i < m; ++i) cj[i] = 0.0;
        } else if (beta != 1.0) {
            for (i64 i = 0; i < m; ++i) cj[i] *= beta;
        }
        for (i64 l = 0; l < k; ++l) {
            double b = (tb == 't') ? B[j + ldb * l] : B[l + ldb * j];
            double s = alpha * b;
            const double *al = ap + ldap * l;
            for (i64 i = 0; i < m; ++i) cj[i] += al[i] * s;
        }
    }
    free(at);
}

/* ------------------------------------------------------------------------ */
/* cube permutations: new(x,y,z) = old(sigma(x,y,z)), n^3 cube, in place     */
/*   id1 (x,z,y)  id2 (y,x,z)  id3 (z,y,x)  id4 (y,z,x)  id5 (z,x,y)          */
/* ------------------------------------------------------------------------ */
#define C3(x, y, z, n) ((x) + (n) * ((y) + (n) * (z)))

static void perm3(double *v, i64 n, int id)
{
    size_t n3 = (size_t)(n * n * n);
    double *old = (double *)malloc(sizeof(double) * n3);
    memcpy(old, v, sizeof(double) * n3);
    for (i64 z = 0; z < n; ++z)
        for (i64 y = 0; y < n; ++y)
            for (i64 x = 0; x < n; ++x) {
                i64 s;
                switch (id) {
                case 1: s = C3(x, z, y, n); break;
                case 2: s = C3(y, x, z, n); break;
                case 3: s = C3(z, y, x, n); break;
                case 4: s = C3(y, z, x, n); break;
                default: s = C3(z, x, y, n); break;
                }
                v[C3(x, y, z, n)] = old[s];
            }
    free(old);
}

/* SYMT311: symmetrise a pair of indices, V <- V + V^(pair swap) (ddcc_aux.f:173-228) */
static void symt311(double *v, i64 n, int id)
{
    int p = (id == 23) ? 1 : (id == 12) ? 2 : 3;
    size_t n3 = (size_t)(n * n * n);
    double *old = (double *)malloc(sizeof(double) * n3);
    memcpy(old, v, sizeof(double) * n3);
    perm3(old, n, p);                        /* old <- swapped copy */
    for (size_t q = 0; q < n3; ++q) v[q] += old[q];
    free(old);
}

static void zerot3(double *v, i64 n)         /* W(a,a,a) = 0 */
{
    for (i64 a = 0; a < n; ++a) v[C3(a, a, a, n)] = 0.0;
}

/* ------------------------------------------------------------------------ */
/* problem data                                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    i64 no, nu;
    const double *eh, *ep, *t1, *t2, *vm, *voe, *vvvo;
} prob_t;

/* T2(:,:,:,x) as an (nu^2, no) block and T2(:,:,y,x) as an (nu,nu) block */
static const double *t2x(const prob_t *p, i64 x) { return p->t2 + p->nu * p->nu * p->no * x; }
static const double *t2yx(const prob_t *p, i64 y, i64 x)
{
    return p->t2 + p->nu * p->nu * (y + p->no * x);
}
/* VM(:,:,x,y) as (no,nu); VOE(:,:,x,y) as (nu,nu) */
static const double *vmxy(const prob_t *p, i64 x, i64 y)
{
    return p->vm + p->no * p->nu * (x + p->no * y);
}
static const double *voexy(const prob_t *p, i64 x, i64 y)
{
    return p->voe + p->nu * p->nu * (x + p->no * y);
}

/* VE_x(d,p,q) = VVVO(d,q,p,x): the fetched patch after unpack + trant3_1 */
static void make_ve(const prob_t *p, i64 x, double *ve)
{
    i64 nu = p->nu;
    const double *s = p->vvvo + nu * nu * nu * x;
    for (i64 q = 0; q < nu; ++q)
        for (i64 pp = 0; pp < nu; ++pp)
            for (i64 d = 0; d < nu; ++d)
                ve[C3(d, pp, q, nu)] = s[C3(d, q, pp, nu)];
}

/* ijk_tuple: W for a distinct tuple (0-based i>j>k) into v3 */
static void ijk_tuple(const prob_t *p, i64 i, i64 j, i64 k,
                      const double *ve_i, const double *ve_j, const double *ve_k, double *v3)
{
    i64 no = p->no, nu = p->nu, nu2 = nu * nu;
    gemm('n', 'n', nu2, nu, no, -1.0, t2x(p, j), nu2, vmxy(p, k, i), no, 0.0, v3, nu2);   /* #2  */
    gemm('t', 't', nu2, nu, nu, 1.0, ve_j, nu, t2yx(p, i, k), nu, 1.0, v3, nu2);          /* #9  */
    perm3(v3, nu, 1);
    gemm('t', 't', nu, nu2, no, -1.0, vmxy(p, j, i), no, t2x(p, k), nu2, 1.0, v3, nu);    /* #4  */
    gemm('n', 'n', nu, nu2, nu, 1.0, t2yx(p, i, j), nu, ve_k, nu, 1.0, v3, nu);           /* #7  */
    perm3(v3, nu, 4);
    gemm('n', 'n', nu2, nu, no, -1.0, t2x(p, i), nu2, vmxy(p, k, j), no, 1.0, v3, nu2);   /* #0  */
    gemm('t', 't', nu2, nu, nu, 1.0, ve_i, nu, t2yx(p, j, k), nu, 1.0, v3, nu2);          /* #11 */
    gemm('n', 'n', nu, nu2, nu, 1.0, t2yx(p, k, i), nu, ve_j, nu, 1.0, v3, nu);           /* #8  */
    gemm('t', 't', nu, nu2, no, -1.0, vmxy(p, i, k), no, t2x(p, j), nu2, 1.0, v3, nu);    /* #3  */
    perm3(v3, nu, 1);
    gemm('n', 'n', nu2, nu, no, -1.0, t2x(p, i), nu2, vmxy(p, j, k), no, 1.0, v3, nu2);   /* #1  */
    gemm('t', 't', nu2, nu, nu, 1.0, ve_i, nu, t2yx(p, k, j), nu, 1.0, v3, nu2);          /* #10 */
    gemm('n', 'n', nu, nu2, nu, 1.0, t2yx(p, j, i), nu, ve_k, nu, 1.0, v3, nu);           /* #6  */
    gemm('t', 't', nu, nu2, no, -1.0, vmxy(p, i, j), no, t2x(p, k), nu2, 1.0, v3, nu);    /* #5  */
    perm3(v3, nu, 1);
    zerot3(v3, nu);
}

/* t1wt3_ijk: E[T] partial and the v1 rows i,j,k for a distinct tuple */
static double t1wt3_ijk(const prob_t *p, i64 i, i64 j, i64 k, const double *v3, double *v1)
{
    i64 nu = p->nu;
    const double *vij = voexy(p, i, j), *vji = voexy(p, j, i), *vik = voexy(p, i, k);
    const double *vki = voexy(p, k, i), *vjk = voexy(p, j, k), *vkj = voexy(p, k, j);
    double dijk = p->eh[i] + p->eh[j] + p->eh[k];
    double x3 = 0.0;
#define V(a, b, c) v3[C3(a, b, c, nu)]
#define M2(m, a, b) m[(a) + nu * (b)]
    for (i64 c = 0; c < nu; ++c)
        for (i64 b = 0; b < nu; ++b)
            for (i64 a = 0; a <= b; ++a) {
                if (a == b && b == c) continue;
                double denom = 1.0 / (dijk - (p->ep[a] + p->ep[b] + p->ep[c]));
                double d1 = V(a, b, c);
                double d2 = V(a, c, b) + V(c, b, a) + V(b, a, c);
                double d3 = V(b, c, a) + V(c, a, b);
                x3 += (8.0 * d1 - 4.0 * d2 + 2.0 * d3) * d1 * denom;
                if (a == b) continue;
                d1 = V(b, a, c);
                d2 = V(b, c, a) + V(c, a, b) + V(a, b, c);
                d3 = V(a, c, b) + V(c, b, a);
                x3 += (8.0 * d1 - 4.0 * d2 + 2.0 * d3) * d1 * denom;
                double z1 = (V(a, b, c) - V(b, a, c)) * 2.0 - V(a, c, b) + V(b, c, a);
                double z2 = (V(a, c, b) - V(b, c, a)) * 2.0 - V(a, b, c) + V(b, a, c);
                double z3 = (V(b, a, c) - V(a, b, c)) * 2.0 - V(c, a, b) + V(c, b, a);
                double z4 = (V(b, c, a) - V(a, c, b)) * 2.0 - V(c, b, a) + V(c, a, b);
                double z5 = (V(c, a, b) - V(c, b, a)) * 2.0 - V(b, a, c) + V(a, b, c);
                double z6 = (V(c, b, a) - V(c, a, b)) * 2.0 - V(b, c, a) + V(a, c, b);
                v1[a + nu * i] += (z1 * M2(vjk, b, c) + z2 * M2(vkj, b, c)) * denom;
                v1[b + nu * i] -= (z1 * M2(vjk, a, c) + z2 * M2(vkj, a, c)) * denom;
                v1[a + nu * j] += (z3 * M2(vik, b, c) + z5 * M2(vki, b, c)) * denom;
                v1[b + nu * j] -= (z3 * M2(vik, a, c) + z5 * M2(vki, a, c)) * denom;
                v1[a + nu * k] += (z4 * M2(vij, b, c) + z6 * M2(vji, b, c)) * denom;
                v1[b + nu * k] -= (z4 * M2(vij, a, c) + z6 * M2(vji, a, c)) * denom;
            }
#undef V
#undef M2
    return (i == j || j == k) ? 0.5 * x3 : x3;
}

/* T3SQUA: E[T] partial of a (degenerate) tuple, full (a,b,c) loop, weight 1/2 */
static double t3squa(const prob_t *p, i64 i, i64 j, i64 k, const double *t3)
{
    i64 nu = p->nu;
    double dijk = p->eh[i] + p->eh[j] + p->eh[k];
    double x3 = 0.0;
#define T(a, b, c) t3[C3(a, b, c, nu)]
    for (i64 c = 0; c < nu; ++c)
        for (i64 b = 0; b < nu; ++b)
            for (i64 a = 0; a < nu; ++a) {
                if (a == b && b == c) continue;
                double denom = 1.0 / (dijk - (p->ep[a] + p->ep[b] + p->ep[c]));
                double d1 = T(a, b, c);
                double d2 = T(a, c, b) + T(c, b, a) + T(b, a, c);
                double d3 = T(b, c, a) + T(c, a, b);
                x3 += (8.0 * d1 - 4.0 * d2 + 2.0 * d3) * d1 * denom;
            }
#undef T
    return (i == j || j == k) ? 0.5 * x3 : x3;
}

static void adt3den(const prob_t *p, double deh, double *t3)
{
    i64 nu = p->nu;
    for (i64 a = 0; a < nu; ++a)
        for (i64 b = 0; b < nu; ++b)
            for (i64 c = 0; c < nu; ++c)
                t3[C3(c, b, a, nu)] /= deh - p->ep[a] - p->ep[b] - p->ep[c];
}

/* T1WT3IJK: Z = SMT3FOUR(TI), then v1(:,i) += Z * VOE(:,:,j,k) (nu x 1 x nu^2 GEMM) */
static void t1wt3ijk(const prob_t *p, i64 i, i64 j, i64 k, const double *ti, double *z, double *v1)
{
    i64 nu = p->nu;
    for (i64 c = 0; c < nu; ++c) {
        for (i64 b = 0; b < nu; ++b) z[C3(b, b, c, nu)] = 0.0;
        for (i64 b = 1; b < nu; ++b)
            for (i64 a = 0; a < b; ++a) {
                double t = (ti[C3(a, b, c, nu)] - ti[C3(b, a, c, nu)]) * 2.0
                           - ti[C3(a, c, b, nu)] + ti[C3(b, c, a, nu)];
                z[C3(a, b, c, nu)] = t;
                z[C3(b, a, c, nu)] = -t;
            }
    }
    gemm('n', 'n', nu, 1, nu * nu, 1.0, z, nu, voexy(p, j, k), nu * nu, 1.0, v1 + nu * i, nu);
}

/* (i,j,j), i>j: ijj_tuple_formv3 + SYMT311(23) + ZEROT3 + T3SQUA + ADT3DEN + DRT1WT3IJ */
static double tuple_ijj(const prob_t *p, i64 i, i64 j, const double *ve_i, const double *ve_j,
                        double *v3, double *t3, double *v1)
{
    i64 no = p->no, nu = p->nu, nu2 = nu * nu;
    gemm('t', 't', nu, nu2, no, -1.0, vmxy(p, i, j), no, t2x(p, j), nu2, 0.0, v3, nu);   /* #2 */
    gemm('n', 'n', nu, nu2, nu, 1.0, t2yx(p, j, i), nu, ve_j, nu, 1.0, v3, nu);          /* #3 */
    perm3(v3, nu, 3);
    gemm('n', 'n', nu, nu2, nu, 1.0, t2yx(p, i, j), nu, ve_j, nu, 1.0, v3, nu);          /* #4 */
    gemm('t', 't', nu, nu2, no, -1.0, vmxy(p, j, i), no, t2x(p, j), nu2, 1.0, v3, nu);   /* #1 */
    perm3(v3, nu, 1);
    gemm('n', 'n', nu, nu2, nu, 1.0, t2yx(p, j, j), nu, ve_i, nu, 1.0, v3, nu);          /* #5 */
    gemm('t', 't', nu, nu2, no, -1.0, vmxy(p, j, j), no, t2x(p, i), nu2, 1.0, v3, nu);   /* #0 */
    perm3(v3, nu, 2);
    symt311(v3, nu, 23);
    zerot3(v3, nu);
    double e = t3squa(p, i, j, j, v3);
    adt3den(p, p->eh[i] + p->eh[j] + p->eh[j], v3);
    t1wt3ijk(p, i, j, j, v3, t3, v1);
    perm3(v3, nu, 2);
    t1wt3ijk(p, j, i, j, v3, t3, v1);
    perm3(v3, nu, 1);
    t1wt3ijk(p, j, j, i, v3, t3, v1);
    return e;
}

/* (i,i,j), i>j: iij_tuple_formv3 + SYMT311(12) + ZEROT3 + T3SQUA + ADT3DEN + DRT1WT3JK */
static double tuple_iij(const prob_t *p, i64 i, i64 j, const double *ve_i, const double *ve_j,
                        double *v3, double *t3, double *v1)
{
    i64 no = p->no, nu = p->nu, nu2 = nu * nu;
    gemm('n', 'n', nu2, nu, no, -1.0, t2x(p, i), nu2, vmxy(p, i, j), no, 0.0, v3, nu2);  /* #1 */
    gemm('n', 'n', nu, nu2, nu, 1.0, t2yx(p, i, i), nu, ve_j, nu, 1.0, v3, nu);          /* #3 */
    perm3(v3, nu, 2);
    gemm('n', 'n', nu2, nu, no, -1.0, t2x(p, j), nu2, vmxy(p, i, i), no, 1.0, v3, nu2);  /* #2 */
    perm3(v3, nu, 1);
    gemm('n', 'n', nu, nu2, nu, 1.0, t2yx(p, i, j), nu, ve_i, nu, 1.0, v3, nu);          /* #5 */
    perm3(v3, nu, 3);
    gemm('n', 'n', nu2, nu, no, -1.0, t2x(p, i), nu2, vmxy(p, j, i), no, 1.0, v3, nu2);  /* #0 */
    gemm('n', 'n', nu, nu2, nu, 1.0, t2yx(p, j, i), nu, ve_i, nu, 1.0, v3, nu);          /* #4 */
    symt311(v3, nu, 12);
    zerot3(v3, nu);
    double e = t3squa(p, i, i, j, v3);
    adt3den(p, p->eh[i] + p->eh[i] + p->eh[j], v3);
    t1wt3ijk(p, i, i, j, v3, t3, v1);
    perm3(v3, nu, 1);
    t1wt3ijk(p, i, j, i, v3, t3, v1);
    perm3(v3, nu, 2);
    t1wt3ijk(p, j, i, i, v3, t3, v1);
    return e;
}

/* DDCC_T_TASK: linear id -> (i>j>k), i outermost (0-based) */
static void task_decode(i64 t, i64 no, i64 *ri, i64 *rj, i64 *rk)
{
    i64 cnt = 0;
    for (i64 i = 0; i < no; ++i)
        for (i64 j = 0; j < i; ++j)
            for (i64 k = 0; k < j; ++k) {
                if (cnt == t) { *ri = i; *rj = j; *rk = k; return; }
                ++cnt;
            }
    *ri = *rj = *rk = -1;
}

/* ------------------------------------------------------------------------ */
/* exported checker entry points (ctypes)                                    */
/* ------------------------------------------------------------------------ */

/* Full driver over distinct tuples [lo,hi) (-1 => all) plus, if degenerate,
 * every (i,j,j) and (i,i,j).  v1 (nu,no) is overwritten; etd/ets returned. */
int port_triples(i64 no, i64 nu, const double *eh, const double *ep, const double *t1,
                 const double *t2, const double *vm, const double *voe, const double *vvvo,
                 i64 lo, i64 hi, int degenerate, double *v1, double *ets, double *etd)
{
    prob_t p = {no, nu, eh, ep, t1, t2, vm, voe, vvvo};
    i64 nu3 = nu * nu * nu;
    i64 ntup = no * (no - 1) * (no - 2) / 6;
    if (lo < 0) lo = 0;
    if (hi < 0 || hi > ntup) hi = ntup;
    double *ve_i = malloc(sizeof(double) * nu3), *ve_j = malloc(sizeof(double) * nu3);
    double *ve_k = malloc(sizeof(double) * nu3), *v3 = malloc(sizeof(double) * nu3);
    double *t3 = malloc(sizeof(double) * nu3);
    if (!ve_i || !ve_j || !ve_k || !v3 || !t3) return -1;
    memset(v1, 0, sizeof(double) * (size_t)(nu * no));
    double e = 0.0;
    for (i64 t = lo; t < hi; ++t) {
        i64 i, j, k;
        task_decode(t, no, &i, &j, &k);
        make_ve(&p, i, ve_i);
        make_ve(&p, j, ve_j);
        make_ve(&p, k, ve_k);
        ijk_tuple(&p, i, j, k, ve_i, ve_j, ve_k, v3);
        e += t1wt3_ijk(&p, i, j, k, v3, v1);
    }
    if (degenerate) {
        for (i64 i = 0; i < no; ++i)
            for (i64 j = 0; j < i; ++j) {
                make_ve(&p, i, ve_i);
                make_ve(&p, j, ve_j);
                e += tuple_ijj(&p, i, j, ve_i, ve_j, v3, t3, v1);
                e += tuple_iij(&p, i, j, ve_i, ve_j, v3, t3, v1);
            }
    }
    double s = 0.0;
    for (i64 i = 0; i < no; ++i)
        for (i64 a = 0; a < nu; ++a) s += t1[i + no * a] * v1[a + nu * i];
    *ets = 2.0 * s;
    *etd = e;
    free(ve_i); free(ve_j); free(ve_k); free(v3); free(t3);
    return 0;
}

/* One tuple: kind 0 = (i,j,k) distinct, 1 = (i,j,j), 2 = (i,i,j); 0-based.
 * v3 receives W as the reference leaves it; v1 is accumulated; etd partial returned. */
int port_tuple(i64 no, i64 nu, const double *eh, const double *ep, const double *t1,
               const double *t2, const double *vm, const double *voe, const double *vvvo,
               int kind, i64 i, i64 j, i64 k, double *v3, double *v1, double *etd)
{
    prob_t p = {no, nu, eh, ep, t1, t2, vm, voe, vvvo};
    i64 nu3 = nu * nu * nu;
    double *ve_i = malloc(sizeof(double) * nu3), *ve_j = malloc(sizeof(double) * nu3);
    double *ve_k = malloc(sizeof(double) * nu3), *t3 = malloc(sizeof(double) * nu3);
    if (!ve_i || !ve_j || !ve_k || !t3) return -1;
    make_ve(&p, i, ve_i);
    make_ve(&p, j, ve_j);
    if (kind == 0) {
        make_ve(&p, k, ve_k);
        ijk_tuple(&p, i, j, k, ve_i, ve_j, ve_k, v3);
        *etd = t1wt3_ijk(&p, i, j, k, v3, v1);
    } else if (kind == 1) {
        *etd = tuple_ijj(&p, i, j, ve_i, ve_j, v3, t3, v1);
    } else {
        *etd = tuple_iij(&p, i, j, ve_i, ve_j, v3, t3, v1);
    }
    free(ve_i); free(ve_j); free(ve_k); free(t3);
    return 0;
}

/* Permutation / symmetrisation of an n^3 cube: id 1..5 perm, 23/12/13 SYMT311 */
void port_cube_op(i64 n, int id, double *v)
{
    if (id >= 1 && id <= 5) perm3(v, n, id);
    else symt311(v, n, id);
}

/* Linear distinct-tuple id -> (i,j,k), 0-based */
void port_task(i64 t, i64 no, i64 *ijk)
{
    task_decode(t, no, &ijk[0], &ijk[1], &ijk[2]);
}
