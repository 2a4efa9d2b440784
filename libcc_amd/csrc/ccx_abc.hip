// Virtual-triple view of the (T) intermediate: for a fixed virtual triple (a,b,c) the
// occupied cube V3(i,j,k) = W_ijk(a,b,c), the quantity the triples-abc prototype
// abc_tuple_formv3 assembles from 12 GEMMs of shape no^2 x nu x no and no x no^2 x no
// (src/triples-abc/abc_tuple.f90:1-81), and E[T] re-summed over virtual triples.
//
// W comes from the closed form on the FP64 operand slabs the context already holds
// (AX[x](a,b,K) = VVVO(K,b,a,x) for K < nu, T2(a,b,K-nu,x) after; BX[y,z](c,K) = T2(c,K,y,z),
// then -VM(K-nu,c,z,y)):
//   W_ijk(a,b,c) = sum over the 6 simultaneous permutations s of the pairs (i,a),(j,b),(k,c)
//                  of sum_K AX[i'](a',b',K) BX[j',k'](c',K).
// This path is a cross-check and a surface for virtual-triple callers, not the hot path:
// one work group per triple, one lane per (i,j,k), FP64 FMA on cached slab reads.

#include <hip/hip_runtime.h>

namespace {

constexpr int ABC_TB = 16;        // virtual block edge of the slab layouts
constexpr int ABC_THR = 256;
__constant__ int ABC_PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// V3 of triple t = (abc[4t], abc[4t+1], abc[4t+2]) into w[t * no^3 ...], column-major
// (i fastest).  V3(i,i,i) is set to 0, as zerot3 does at the end of abc_tuple_formv3
// (src/triples-abc/abc_tuple.f90:77); E[T] does not depend on those elements.
__global__ void __launch_bounds__(ABC_THR) w_abc_kernel(const double *__restrict__ axr, const double *__restrict__ bx,
                                                        const int *__restrict__ trip, int ntrip, int no, int nu,
                                                        int nb, int kp, double *__restrict__ w)
{
    const int t = blockIdx.x;
    if (t >= ntrip) return;
    const int v[3] = {trip[4 * t], trip[4 * t + 1], trip[4 * t + 2]};
    const int no3 = no * no * no;
    const int kn = nu + no;
    for (int e = threadIdx.x; e < no3; e += ABC_THR) {
        const int o[3] = {e % no, (e / no) % no, e / (no * no)};
        double acc = 0.0;
        if (!(o[0] == o[1] && o[1] == o[2])) {
            for (int s = 0; s < 6; ++s) {
                const int p0 = ABC_PERM[s][0], p1 = ABC_PERM[s][1], p2 = ABC_PERM[s][2];
                const int x = o[p0], y = o[p1], z = o[p2];
                const int av = v[p0], bv = v[p1], cv = v[p2];
                // row layout of AX (pack_ax_kernel<0>): panel (x, a-block, b-block), then
                // [alpha/2][K/4][lane = (K%4)*16 + beta][alpha%2]
                const int al = av % ABC_TB, be = bv % ABC_TB;
                const double *ap = axr + (((long long)x * nb + av / ABC_TB) * nb + bv / ABC_TB) * (256LL * kp)
                                   + (long long)(al >> 1) * (kp / 4) * 128 + be * 2 + (al & 1);
                // BX (pack_bx_kernel): [y][z][c-block][K][gamma]
                const double *bp = bx + (((long long)y * no + z) * nb + cv / ABC_TB) * ((long long)kp * 16) + cv % ABC_TB;
                double sum = 0.0;
                for (int K = 0; K < kn; ++K) sum = fma(ap[(K >> 2) * 128 + (K & 3) * 32], bp[(long long)K * 16], sum);
                acc += sum;
            }
        }
        w[(long long)t * no3 + e] = acc;
    }
}

// X_abc = sum_ijk V(ijk) [8 V(ijk) - 4 (V(ikj) + V(kji) + V(jik)) + 2 (V(kij) + V(jki))] / D
// with D = e_i + e_j + e_k - e_a - e_b - e_c, times the triple's weight (1 for a > b > c,
// 1/2 for a > b = c and a = b > c).  It is the ijk-driven t1wt3_ijk / T3SQUA sum
// (src/triples-ijk/t1wt3.f90:53-69, src/triples-ijk/t3squa_acc.f:18-56) with the roles of
// the occupied and the virtual indices exchanged through the pair-permutation symmetry of W.
// Fixed-order reduction: each lane sums its elements in order, then a tree over the group.
__global__ void __launch_bounds__(ABC_THR) x_abc_kernel(const double *__restrict__ w, const int *__restrict__ trip,
                                                        int ntrip, int no, const double *__restrict__ eh,
                                                        const double *__restrict__ ep, double *__restrict__ x)
{
    __shared__ double part[ABC_THR];
    const int t = blockIdx.x;
    if (t >= ntrip) return;
    const double eabc = ep[trip[4 * t]] + ep[trip[4 * t + 1]] + ep[trip[4 * t + 2]];
    const double wt = trip[4 * t + 3] == 1 ? 1.0 : 0.5;
    const int no3 = no * no * no;
    const double *V = w + (long long)t * no3;
    double s = 0.0;
    for (int e = threadIdx.x; e < no3; e += ABC_THR) {
        const int i = e % no, j = (e / no) % no, k = e / (no * no);
        auto at = [&](int p, int q, int r) { return V[p + no * (q + no * r)]; };
        const double vijk = V[e];
        const double br = 8.0 * vijk - 4.0 * (at(i, k, j) + at(k, j, i) + at(j, i, k)) + 2.0 * (at(k, i, j) + at(j, k, i));
        s += vijk * br / (eh[i] + eh[j] + eh[k] - eabc);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = ABC_THR / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) x[t] = part[0] * wt;
}

}  // namespace

extern "C" {

int ccx_hip_w_abc(const void *axr, const void *bx, const int *trip, int ntrip, int no, int nu, int nb, int kp,
                  double *w, void *stream)
{
    if (ntrip <= 0) return 0;
    if (kp % 4 != 0 || kp < nu + no) return 3;
    w_abc_kernel<<<ntrip, ABC_THR, 0, (hipStream_t)stream>>>((const double *)axr, (const double *)bx, trip, ntrip,
                                                             no, nu, nb, kp, w);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_x_abc(const double *w, const int *trip, int ntrip, int no, const double *eh, const double *ep,
                  double *x, void *stream)
{
    if (ntrip <= 0) return 0;
    x_abc_kernel<<<ntrip, ABC_THR, 0, (hipStream_t)stream>>>(w, trip, ntrip, no, eh, ep, x);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // extern "C"
