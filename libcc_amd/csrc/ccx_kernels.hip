// Closed-shell (T) triples correction on gfx950: device kernels and their thin C ABI.
//
// Formulation (closed form of what src/triples-ijk/ijk_tuple.f90:56-84 builds with 12
// DGEMMs and trant3 steps):
//
//   W_ijk(a,b,c) = sum over the 6 simultaneous permutations s of the pairs (i,a),(j,b),(k,c)
//                  of  sum_K AX[x](alpha,beta,K) * BX[y,z](gamma,K)
//   AX[x](alpha,beta,K) = VVVO(K,beta,alpha,x)      K <  nu
//                       = T2(alpha,beta,K-nu,x)     nu <= K < nu+no
//   BX[y,z](gamma,K)    = T2(gamma,K,y,z)           K <  nu
//                       = -VM(K-nu,gamma,z,y)       nu <= K < nu+no
//
// with (x,alpha),(y,beta),(z,gamma) the permuted pairs.  Particle (VVVO*T2, K=nu) and
// hole (T2*VM, K=no) contractions are one GEMM over K = nu+no.
//
// Work item = (occupied tuple, virtual block orbit).  The virtual cube is cut into
// 16^3 blocks; a work item owns the blocks (A<=B<=C) and accumulates, for the
// orderings P of its occupied tuple, w_P(a,b,c) = W_P(a,b,c) over (a,b,c) in A x B x C.
// Because W_{pi.P}(pi.e) = W_P(e) (pi permuting tuple positions and virtual positions
// alike), the six w_P hold W_ijk at the six positional permutations of (a,b,c), in the
// SAME register slot, so the energy needs no transposition of W:
//   sum_s f(s.e) = Q(w) / D,  Q = 6 sum w_P^2 + 2 E^2 + 2 O^2 - 8 E O,
//   E = w over even orderings, O = w over odd orderings,
// f being the summand of t1wt3_ijk's E[T] loop (src/triples-ijk/t1wt3.f90:143-159).
// Degenerate tuples (i,j,j)/(i,i,j) have three distinct orderings; blocks with equal
// edges (A=B, B=C, A=B=C) get the missing orderings by an in-block transpose, so no
// ordering is computed twice.  The index permutations (trant3,
// src/common/common_cc.f90:62-333) become the choice of operand slab and orientation.
//
// FP64 MFMA v_mfma_f64_16x16x4_f64: lane l supplies A[l%16][l/16] and B[l/16][l%16]
// and holds D[l/16 + 4v][l%16], v = 0..3 (checked on the device).

#include "ccx_item.h"
#include <stdlib.h>
#include <string.h>


// ---------------------------------------------------------------------------------
// packing kernels (once per problem)
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double ax_value(const double *__restrict__ vvvo, const double *__restrict__ t2,
                                           int no, int nu, int x, int ag, int bg, long long K)
{
    if (ag >= nu || bg >= nu) return 0.0;
    if (K < nu) return vvvo[K + (long long)nu * (bg + (long long)nu * (ag + (long long)nu * x))];
    if (K < nu + no) return t2[ag + (long long)nu * (bg + (long long)nu * ((K - nu) + (long long)no * x))];
    return 0.0;
}

// COLS = 0: row layout [..][alpha][kb][k4][beta]; COLS = 1: column layout [..][beta][kb][k4][alpha].
// Packs the panels of nx occupied indices x = 0..nx-1 from vvvo + x nu^3 and t2 + x nu^2 no (nx = no
// for the whole slab; 1 for one index, with the pointers at that index).
template <int COLS, class T>
__global__ void pack_ax_kernel(const double *__restrict__ vvvo, const double *__restrict__ t2,
                               T *__restrict__ ax, int no, int nu, int nb, int kp, int nx)
{
    const long long per_panel = 256LL * kp;
    const long long total = (long long)nx * nb * nb * per_panel;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long r = idx % per_panel;
        long long pnl = idx / per_panel;
        // [outer/2][kp/4][64 lanes][2]: the fragments of the two batches (outer indices 2w,
        // 2w+1) a wave owns interleave per lane, so one 2-wide load fetches both
        const int bbit = (int)(r & 1);
        const int lane = (int)((r >> 1) & 63);
        const long long rr = r >> 7;
        const long long K = (rr % (kp / 4)) * 4 + (lane >> 4);   // kb*4 + k4
        const int outer = (int)(rr / (kp / 4)) * 2 + bbit;       // alpha (rows) / beta (columns)
        const int inner = lane & 15;                              // beta (rows) / alpha (columns)
        const int q = pnl % nb; pnl /= nb;
        const int pblk = pnl % nb; pnl /= nb;
        const int x = (int)pnl;
        const int al = COLS ? inner : outer, be = COLS ? outer : inner;
        ax[idx] = (T)ax_value(vvvo, t2, no, nu, x, pblk * TB + al, q * TB + be, K);
    }
}

template <class T>
__global__ void pack_bx_kernel(const double *__restrict__ t2, const double *__restrict__ vm,
                               T *__restrict__ bx, int no, int nu, int nb, int kp)
{
    const long long total = (long long)no * no * nb * kp * 16;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long r = idx;
        const int gl = r & 15; r >>= 4;
        const int K = r % kp; r /= kp;
        const int gb = r % nb; r /= nb;
        const int z = r % no; r /= no;
        const int y = (int)r;
        const int g = gb * TB + gl;
        double v = 0.0;
        if (g < nu) {
            if (K < nu)
                v = t2[g + (long long)nu * (K + (long long)nu * (y + (long long)no * z))];
            else if (K < nu + no)
                v = -vm[(K - nu) + (long long)no * (g + (long long)nu * (z + (long long)no * y))];
        }
        bx[idx] = (T)v;
    }
}

// BX of one ordered pair (y,z): BX[y,z](g,K) = T2(g,K,y,z) for K < nu, -VM(K-nu,g,z,y) for K < nu+no;
// t2yz = T2(:,:,y,z) (nu^2), vmzy = VM(:,:,z,y) (no*nu), bx = the pair's [nb][kp][16] panel set.
__global__ void pack_bx_pair_kernel(const double *__restrict__ t2yz, const double *__restrict__ vmzy,
                                    double *__restrict__ bx, int no, int nu, int nb, int kp)
{
    const long long total = (long long)nb * kp * 16;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int gl = (int)(idx & 15);
        const int K = (int)((idx >> 4) % kp);
        const int g = (int)((idx >> 4) / kp) * TB + gl;
        double v = 0.0;
        if (g < nu) {
            if (K < nu) v = t2yz[g + (long long)nu * K];
            else if (K < nu + no) v = -vmzy[(K - nu) + (long long)no * g];
        }
        bx[idx] = v;
    }
}

// VVVO(d,q,p,x) = VE_x(d,p,q): the (2,3) index swap of one occupied index's cube
__global__ void ve_to_vvvo_kernel(const double *__restrict__ ve, double *__restrict__ vvvo, int nu)
{
    const long long n3 = (long long)nu * nu * nu;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n3;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long d = idx % nu, q = (idx / nu) % nu, p = idx / ((long long)nu * nu);
        vvvo[idx] = ve[d + nu * (p + (long long)nu * q)];
    }
}

__global__ void pack_voe_kernel(const double *__restrict__ voe, double *__restrict__ out, int no,
                                int nu, int nup)
{
    const long long total = (long long)no * no * nup * nup;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long r = idx;
        const int y = r % nup; r /= nup;
        const int x = r % nup; r /= nup;
        const int rr = r % no; r /= no;
        const int q = (int)r;
        double v = 0.0;
        if (x < nu && y < nu)
            v = voe[x + (long long)nu * (y + (long long)nu * (q + (long long)no * rr))];
        out[idx] = v;
    }
}

// Distributed (nutr, nu*no) layout -> VVVO(nu,nu,nu,no): column c + nu*x holds the
// lower-packed symmetric (p,q) pair block, row p*(p+1)/2 + q for q <= p (0-based),
// as DDCC_T_GETVE / CPYTSQ expand it (src/triples-ijk/common_functions.f:43-60,106-125).
__global__ void unpack_vvvo_kernel(const double *__restrict__ packed, double *__restrict__ full, int no, int nu)
{
    const long long nutr = (long long)nu * (nu + 1) / 2;
    const long long total = (long long)nu * nu * nu * no;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int p = idx % nu;
        const int q = (idx / nu) % nu;
        const long long col = idx / ((long long)nu * nu);   // c + nu*x
        const int hi = p > q ? p : q, lo = p > q ? q : p;
        full[idx] = packed[(long long)hi * (hi + 1) / 2 + lo + nutr * col];
    }
}

// ---------------------------------------------------------------------------------
// split forms: fixed-point digit slices of the FP64 operand slabs
// ---------------------------------------------------------------------------------

// max |x| over a slab (non-negative doubles order like their bit patterns)
__global__ void absmax_kernel(const double *__restrict__ x, long long n, unsigned long long *__restrict__ out)
{
    double m = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[i]));
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// OCP e4m3 byte of an integer |d| <= 16 (every such integer has at most 4 significant bits)
__device__ __forceinline__ unsigned char e4m3_of_int(int d)
{
    if (d == 0) return 0;
    const unsigned char sgn = d < 0 ? 0x80 : 0;
    const int m = d < 0 ? -d : d;
    const int p = 31 - __builtin_clz(m);          // 2^p <= m < 2^(p+1)
    const int frac = ((m << 3) >> p) & 7;         // the three bits after the leading one
    return sgn | (unsigned char)((p + 7) << 3) | (unsigned char)frac;
}

// Slices of a slab in the fragment order of the form's MFMA (see run_phase_split).  One thread
// per group of EL elements of one row (AX: row = the panel's outer index; BX: no outer index),
// inner index lane & 15.  Output: [panel][outer][kb][slice][16-byte unit f][lane][16 B].  The MFMA
// sums over every K of a block, so any placement of K inside a block that A and B share gives
// the same products:
//   bf16 (EL 8): the group is lane's fragment, K = 32 kb + 8 (lane >> 4) + e;
//   fp8 (EL 32): the group is scale block b = lane >> 4 of the 128-K block (K = 128 kb + 32 b + q,
//     q < 32), which v_mfma_scale_f32_16x16x128_f8f6f4 reads as bytes 16 (b >> 1) + (q & 15) of
//     lanes 16 (2 (b & 1) + (q >> 4)) + row and scales with the word of lane 16 b + row (the
//     layout measured on the device); the group's exponent word goes to [panel][outer][kb][lane].
// Source: the FP64 slab, AX [panel][outer/2][K/4][lane 64][2] or BX [panel][K][16 gamma].
// inv_scale = 1 / sigma of the slab.
template <int FK, int BX>
__global__ void pack_split_kernel(const double *__restrict__ src, unsigned char *__restrict__ dst,
                                  int *__restrict__ dexp, long long npanel, int kp, int kps, double inv_scale)
{
    typedef SplitCfg<FK> CF;
    constexpr int NS = CF::NS, F16 = CF::F16, KB = CF::KB;
    constexpr int EL = CF::FMT == 0 ? 8 : 32, EB = CF::FMT == 0 ? 2 : 1;
    const int nkb = kps / KB;
    const long long ngroups = npanel * (BX ? 1 : 16) * nkb * 64;
    for (long long gi = blockIdx.x * (long long)blockDim.x + threadIdx.x; gi < ngroups;
         gi += (long long)gridDim.x * blockDim.x) {
        const int lane = (int)(gi & 63);
        const long long rr = gi >> 6;      // (panel, outer, kb) or (panel, kb)
        const int kb = (int)(rr % nkb);
        const long long po = rr / nkb;     // panel * 16 + outer, or panel
        const int inner = lane & 15;
        double v[EL];
        double m = 0.0;
#pragma unroll
        for (int e = 0; e < EL; ++e) {
            const int K = kb * KB + (lane >> 4) * EL + e;
            double x = 0.0;
            if (K < kp) {
                if (BX) {
                    x = src[po * (16LL * kp) + (long long)K * 16 + inner];
                } else {
                    const int outer = (int)(po % 16);
                    const long long panel = po / 16;
                    x = src[panel * (256LL * kp) + (((long long)(outer >> 1) * (kp / 4) + K / 4) * 64 + (K & 3) * 16 + inner) * 2 +
                            (outer & 1)];
                }
            }
            v[e] = x * inv_scale;          // |v| <= 1
            m = fmax(m, fabs(v[e]));
        }
        int gexp = 0;                      // FP8 form: group exponent, |v| <= 2^gexp
        if constexpr (CF::MX) {
            if (m > 0.0) frexp(m, &gexp);
            if (gexp < -100) gexp = -100;  // a group this small is carried as zero digits
        }
#pragma unroll
        for (int e = 0; e < EL; ++e) {
            int f = (e * EB) / 16, byte = (e * EB) % 16, ol = lane;
            if constexpr (CF::MX) {   // element q = e of scale block b = lane >> 4
                const int b = lane >> 4;
                ol = 16 * (2 * (b & 1) + (e >> 4)) + inner;
                f = b >> 1;
                byte = e & 15;
            }
            double t = v[e];
            if constexpr (CF::MX) t = ldexp(t, -gexp);
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                const long long o = (((rr * NS + sl) * F16 + f) * 64 + ol) * 16 + byte;
                if constexpr (FK == 2) {          // per element: bf16 rounding of what is left
                    const __bf16 h = (__bf16)(float)t;
                    t -= (double)(float)h;
                    *reinterpret_cast<__bf16 *>(dst + o) = h;
                } else if constexpr (FK == 4) {   // fixed point: digits base 2^8
                    t *= 256.0;
                    const double d = rint(t);
                    t -= d;
                    *reinterpret_cast<__bf16 *>(dst + o) = (__bf16)(float)ldexp(d, -8 * (sl + 1));   // exact
                } else {                          // digits base 2^4 as e4m3 integers
                    t *= 16.0;
                    const double d = rint(t);
                    t -= d;
                    dst[o] = gexp > -100 ? e4m3_of_int((int)d) : 0;
                }
            }
        }
        if constexpr (CF::MX) {
            // byte s: biased E8M0 exponent of slice s, 2^(gexp - 4(s+1))
            unsigned int w = 0;
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) w |= (unsigned int)(127 + gexp - 4 * (sl + 1)) << (8 * sl);
            dexp[rr * 64 + lane] = (int)w;
        }
    }
}

// ---------------------------------------------------------------------------------
// persistent (tuple, block-orbit) kernel, FP64 MFMA
// ---------------------------------------------------------------------------------
namespace {

// FK: 0 FP64 slabs, 1 FP32 slabs (16 K per fragment load), 2-4 split slices (SplitCfg).  Always
// inlined: as a called function every item paid a call frame and callee-saved register
// spills in scratch memory (0.06 TB of writes per launch at no=24, nu=216).
template <int MODE, class T, int FK>
__device__ __forceinline__ void do_item(const TriplesArgs &g, double *lds, double (*red)[112], const ItemCtx &it, long long item,
                        int tkind, int mult, bool swap_bc, int wave, int lane, int tid)
{
    typename Mfma<T>::acc_t acc[6][BPW];
    f4 acc1[FK == 4 ? 6 : 1][BPW];   // two-accumulator split form: the levels above 0
    if constexpr (FK >= 2) item_w_split<MODE, FK>(g, lds, it, acc, acc1, wave, lane);
    else if constexpr (FK == 1) item_w_f32v<MODE>(g, lds, it, acc, wave, lane);
    else item_w<MODE, T>(g, lds, it, acc, wave, lane);

    const int A = it.blk[0], B = it.blk[1], C = it.blk[2];
    // ---- stage the 36 VOE sub-blocks: [ordered tuple-position pair][ordered block pair]
    for (int idx = tid; idx < 36 * 16 * 4; idx += NTHR) {
        const int quad = idx & 3, row = (idx >> 2) & 15, sb = idx >> 6;
        const int qr = sb / 6, bpair = sb % 6;
        const int y = qr >> 1, z = (qr & 1) ? (y == 2 ? 1 : 2) : (y == 0 ? 1 : 0);
        const int ry = bpair >> 1, rz = (bpair & 1) ? (ry == 2 ? 1 : 2) : (ry == 0 ? 1 : 0);
        const double *src = g.voe + (((long long)pick3(it.tpos, y) * g.no + pick3(it.tpos, z)) * g.nup
                                     + pick3(it.blk, ry) * TB + row) * g.nup + pick3(it.blk, rz) * TB + quad * 4;
        double *d = lds + sb * VBLK + row * VSTR + quad * 4;
        d[0] = src[0]; d[1] = src[1]; d[2] = src[2]; d[3] = src[3];
    }
    __syncthreads();

    // ---- epilogue: energy, v1 partials, optional W store
    const double dijk = g.eh[it.tpos[0]] + g.eh[it.tpos[1]] + g.eh[it.tpos[2]];
    const double wgt = (tkind == 1 ? 1.0 : 0.5) / (double)mult;
    double esum = 0.0;
    double va[4][3], vb[BPW][3], vc[3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int p = 0; p < 3; ++p) va[i][p] = 0.0;
#pragma unroll
    for (int i = 0; i < BPW; ++i)
#pragma unroll
        for (int p = 0; p < 3; ++p) vb[i][p] = 0.0;
    vc[0] = vc[1] = vc[2] = 0.0;
    const int cl = lane & 15;
    const int cg = C * TB + cl;
#pragma unroll
    for (int bb = 0; bb < BPW; ++bb) {
        const int bl = wave * BPW + bb;
        const int bg = B * TB + bl;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int al = Mfma<T>::row(lane, v);
            const int ag = A * TB + al;
            if (ag >= g.nu || bg >= g.nu || cg >= g.nu) continue;
            double w[6];
#pragma unroll
            for (int p = 0; p < 6; ++p) {
                if constexpr (FK == 4) w[p] = ((double)acc[p][bb][v] + (double)acc1[p][bb][v]) * g.wscale;
                else if constexpr (FK >= 2) w[p] = (double)acc[p][bb][v] * g.wscale;
                else w[p] = (double)acc[p][bb][v];
            }
            const double inv = 1.0 / (dijk - g.ep[ag] - g.ep[bg] - g.ep[cg]);
            const double ev = w[0] + w[3] + w[4], od = w[1] + w[2] + w[5];
            const double sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2] + w[3] * w[3] + w[4] * w[4] + w[5] * w[5];
            esum += (6.0 * sq + 2.0 * ev * ev + 2.0 * od * od - 8.0 * ev * od) * inv;
            const int el[3] = {al, bl, cl};
            static_for<6>([&](auto S) __attribute__((always_inline)) {
                constexpr int s = decltype(S)::value;
                constexpr int r0 = PERM[s][0], r1 = PERM[s][1], r2 = PERM[s][2];
                constexpr int bpair = pair_index(r1, r2);
                static_for<6>([&](auto PP) __attribute__((always_inline)) {
                    constexpr int p = decltype(PP)::value;
                    constexpr int q0 = wsrc_comp(p, s, 0);   // sigma
                    constexpr int q1 = wsrc_comp(p, s, 2);   // sigma o (1,0,2)
                    constexpr int q2 = wsrc_comp(p, s, 1);   // sigma o (0,2,1)
                    constexpr int q3 = wsrc_comp(p, s, 3);   // sigma o (1,2,0)
                    constexpr int qr = pair_index(PERM[p][1], PERM[p][2]);
                    const double z = 2.0 * w[q0] - 2.0 * w[q1] - w[q2] + w[q3];
                    const double vv = lds[(qr * 6 + bpair) * VBLK + el[r1] * VSTR + el[r2]];
                    const double c = z * vv * inv;
                    if constexpr (r0 == 0) va[v][PERM[p][0]] += c;
                    else if constexpr (r0 == 1) vb[bb][PERM[p][0]] += c;
                    else vc[PERM[p][0]] += c;
                });
            });
            if (g.wout != nullptr) {
                const int xg[3] = {ag, bg, cg};
                static_for<6>([&](auto S) __attribute__((always_inline)) {
                    constexpr int s = decltype(S)::value;
                    constexpr int q0 = wsrc_comp(0, s, 0);
                    g.wout[xg[PERM[s][0]] + (long long)g.nu * (xg[PERM[s][1]] + (long long)g.nu * xg[PERM[s][2]])] = w[q0];
                });
            }
        }
    }

    // ---- deterministic reductions
    double *rw = red[wave];
#pragma unroll
    for (int v = 0; v < 4; ++v)          // role a: rows (lane>>4)+4v
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const double s = group16_sum(va[v][p]);
            if ((lane & 15) == 0) rw[Mfma<T>::row(lane, v) * 3 + p] = s;
        }
#pragma unroll
    for (int bb = 0; bb < BPW; ++bb)     // role b: rows wave*BPW+bb
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const double s = wave_sum(vb[bb][p]);
            if (lane == 0) rw[48 + bb * 3 + p] = s;
        }
#pragma unroll
    for (int p = 0; p < 3; ++p) {        // role c: rows lane&15
        double s = vc[p];
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (lane < 16) rw[54 + lane * 3 + p] = s;
    }
    {
        const double s = wave_sum(esum);
        if (lane == 0) rw[102] = s;
    }
    __syncthreads();
    double *vp = g.v1_part + item * 144;
    const int slot_b = swap_bc ? 96 : 48, slot_c = swap_bc ? 48 : 96;   // table role order
    if (tid < 48) {
        double s = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < NWAVE; ++w2) s += red[w2][tid];
        vp[tid] = s * wgt;
    } else if (tid < 96) {
        const int r = tid - 48, row = r / 3, p = r % 3;
        vp[slot_b + r] = red[row / BPW][48 + (row % BPW) * 3 + p] * wgt;
    } else if (tid < 144) {
        const int r = tid - 96;
        double s = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < NWAVE; ++w2) s += red[w2][54 + r];
        vp[slot_c + r] = s * wgt;
    } else if (tid == 144) {
        double s = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < NWAVE; ++w2) s += red[w2][102];
        g.e_part[item] = s * wgt;
    }
    __syncthreads();
}

}  // namespace

template <class T, int FK>
__global__ void __launch_bounds__(NTHR, 1) triples_kernel(TriplesArgs g)
{
    extern __shared__ double lds[];
    __shared__ double red[NWAVE][112];
    __shared__ long long next_item;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    for (;;) {
        // Work order: block triple major (default), so the tuples of one block triple run
        // together and the operand panels they share stay in the last-level cache (34% less
        // L2 miss traffic than tuple major at no=24, nu=216); tuple major on request.
        if (tid == 0) {
            const unsigned int p = atomicAdd(g.counter, 1u);
            long long it = -1;
            if (p < g.nitems) it = g.block_major ? (long long)p : (long long)(p % g.nblk) * g.ntup + p / g.nblk;
            next_item = it;
        }
        __syncthreads();
        const long long pos = next_item;   // = q * ntup + t
        __syncthreads();
        if (pos < 0) break;
        const int q = __builtin_amdgcn_readfirstlane((int)(pos / g.ntup));
        const int t = __builtin_amdgcn_readfirstlane((int)(pos % g.ntup));
        const long long item = (long long)t * g.nblk + q;   // storage index
        int tkind, mult;
        ItemCtx it = make_item_ctx(g.tuples, g.blocks, t, q, tkind, mult);
        const bool swap_bc = ragged_block_to_b(g, it, mult);
        if constexpr (pair_modes(FK)) {
            dispatch_mode(tkind, mult, it, [&](auto M) __attribute__((always_inline)) {
                do_item<decltype(M)::value, T, FK>(g, lds, red, it, item, tkind, mult, swap_bc, wave, lane, tid);
            });
        } else {
            if (tkind == 2) do_item<M_IJJ, T, FK>(g, lds, red, it, item, tkind, mult, swap_bc, wave, lane, tid);
            else if (tkind == 3) do_item<M_IIJ, T, FK>(g, lds, red, it, item, tkind, mult, swap_bc, wave, lane, tid);
            else if (mult == 1) do_item<M_FULL, T, FK>(g, lds, red, it, item, tkind, mult, swap_bc, wave, lane, tid);
            else if (mult == 6) do_item<M_AAA, T, FK>(g, lds, red, it, item, tkind, mult, swap_bc, wave, lane, tid);
            else if (it.blk[0] == it.blk[1]) do_item<M_AAC, T, FK>(g, lds, red, it, item, tkind, mult, swap_bc, wave, lane, tid);
            else do_item<M_ACC, T, FK>(g, lds, red, it, item, tkind, mult, swap_bc, wave, lane, tid);
        }
    }
}

// ---------------------------------------------------------------------------------
// finalize: deterministic reductions
// ---------------------------------------------------------------------------------
// v1t[t][a][p] = sum over the blocks of tuple t whose role-r block contains a
__global__ void reduce_v1_tuple_kernel(const double *__restrict__ v1_part, const int *__restrict__ blocks,
                                       double *__restrict__ v1t, int ntup, int nblk, int nu)
{
    const long long total = (long long)ntup * nu * 3;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int p = idx % 3;
        const int a = (idx / 3) % nu;
        const long long t = idx / (3LL * nu);
        const int ab = a / TB, al = a % TB;
        double s = 0.0;
        for (int q = 0; q < nblk; ++q) {
            const int4 bp = reinterpret_cast<const int4 *>(blocks)[q];
            const double *vp = v1_part + (t * nblk + q) * 144;
            if (bp.x == ab) s += vp[al * 3 + p];
            if (bp.y == ab) s += vp[48 + al * 3 + p];
            if (bp.z == ab) s += vp[96 + al * 3 + p];
        }
        v1t[idx] = s;
    }
}

// v1(a,o) (column-major nu x no) = sum over tuples and positions p with tuple[p] == o
__global__ void reduce_v1_kernel(const double *__restrict__ v1t, const int *__restrict__ tuples,
                                 double *__restrict__ v1, int ntup, int nu, int no)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nu * no) return;
    const int a = idx % nu, o = idx / nu;
    double s = 0.0;
    for (int t = 0; t < ntup; ++t) {
        const int4 tp = reinterpret_cast<const int4 *>(tuples)[t];
        const double *r = v1t + ((long long)t * nu + a) * 3;
        if (tp.x == o) s += r[0];
        if (tp.y == o) s += r[1];
        if (tp.z == o) s += r[2];
    }
    v1[idx] = s;
}

// out[0] = etd = sum e_part (compensated, fixed order); out[1] = ets = 2 sum T1(o,a) v1(a,o)
__global__ void __launch_bounds__(1024) finalize_energy_kernel(const double *__restrict__ e_part, long long n,
                                                               const double *__restrict__ t1,
                                                               const double *__restrict__ v1, int no, int nu,
                                                               double *__restrict__ out)
{
    __shared__ double sh[2][1024];
    const int tid = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (long long i = tid; i < n; i += 1024) {
        const double y = e_part[i];
        const double tt = s + y;
        c += (fabs(s) >= fabs(y)) ? ((s - tt) + y) : ((y - tt) + s);
        s = tt;
    }
    double s2 = 0.0, c2 = 0.0;
    for (int i = tid; i < no * nu; i += 1024) {
        const int a = i % nu, o = i / nu;
        const double y = t1[o + no * a] * v1[i];
        const double tt = s2 + y;
        c2 += (fabs(s2) >= fabs(y)) ? ((s2 - tt) + y) : ((y - tt) + s2);
        s2 = tt;
    }
    sh[0][tid] = s + c;
    sh[1][tid] = s2 + c2;
    __syncthreads();
    for (int w = 512; w >= 1; w >>= 1) {
        if (tid < w) {
            sh[0][tid] += sh[0][tid + w];
            sh[1][tid] += sh[1][tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = sh[0][0];
        out[1] = 2.0 * sh[1][0];
    }
}

// E[T] and v1 of one tuple from a given W_ijk cube (nu^3, column-major), the computation of
// t1wt3_ijk (src/triples-ijk/t1wt3.f90:53-102) as a full sum over (b,c) per work group a:
//   x3[a]      = sum_{b,c, not a=b=c} W(abc) (8W(abc) - 4(W(acb)+W(cba)+W(bac)) + 2(W(bca)+W(cab))) / D
//   v1p[a][0]  = sum_{b,c} (Z_ijk(abc) VOE_jk(bc) + Z_ikj(abc) VOE_kj(bc)) / D     (row i)
//   v1p[a][1]  = sum_{b,c} (Z_jik(abc) VOE_ik(bc) + Z_kij(abc) VOE_ki(bc)) / D     (row j)
//   v1p[a][2]  = sum_{b,c} (Z_jki(abc) VOE_ij(bc) + Z_kji(abc) VOE_ji(bc)) / D     (row k)
// with Z_P the antisymmetric combination 2(W_P(abc)-W_P(bac)) - W_P(acb) + W_P(bca) of the
// cube read in ordering P.  voe6 = VOE(:,:,x,y) for xy = ij, ji, ik, ki, jk, kj.  Fixed-order
// tree reduction per work group: deterministic.
__global__ void __launch_bounds__(256) t1wt3_kernel(const double *__restrict__ w, const double *__restrict__ voe6,
                                                    double dijk, const double *__restrict__ ep, int nu,
                                                    double *__restrict__ x3, double *__restrict__ v1p)
{
    __shared__ double sh[4][256];
    const int a = blockIdx.x, tid = threadIdx.x;
    const long long n2 = (long long)nu * nu;
    auto W = [&](int x, int y, int z) { return w[x + (long long)nu * (y + (long long)nu * z)]; };
    double e = 0.0, ri = 0.0, rj = 0.0, rk = 0.0;
    for (long long bc = tid; bc < n2; bc += 256) {
        const int b = (int)(bc % nu), c = (int)(bc / nu);
        if (a == b && b == c) continue;
        const double inv = 1.0 / (dijk - ep[a] - ep[b] - ep[c]);
        const double abc = W(a, b, c), acb = W(a, c, b), bac = W(b, a, c);
        const double bca = W(b, c, a), cab = W(c, a, b), cba = W(c, b, a);
        e += abc * (8.0 * abc - 4.0 * (acb + cba + bac) + 2.0 * (bca + cab)) * inv;
        const double z1 = 2.0 * (abc - bac) - acb + bca;   // ijk
        const double z2 = 2.0 * (acb - bca) - abc + bac;   // ikj
        const double z3 = 2.0 * (bac - abc) - cab + cba;   // jik
        const double z4 = 2.0 * (bca - acb) - cba + cab;   // jki
        const double z5 = 2.0 * (cab - cba) - bac + abc;   // kij
        const double z6 = 2.0 * (cba - cab) - bca + acb;   // kji
        const long long o = b + (long long)nu * c;
        ri += (z1 * voe6[4 * n2 + o] + z2 * voe6[5 * n2 + o]) * inv;
        rj += (z3 * voe6[2 * n2 + o] + z5 * voe6[3 * n2 + o]) * inv;
        rk += (z4 * voe6[0 * n2 + o] + z6 * voe6[1 * n2 + o]) * inv;
    }
    sh[0][tid] = e; sh[1][tid] = ri; sh[2][tid] = rj; sh[3][tid] = rk;
    __syncthreads();
    for (int s2 = 128; s2 >= 1; s2 >>= 1) {
        if (tid < s2)
            for (int q = 0; q < 4; ++q) sh[q][tid] += sh[q][tid + s2];
        __syncthreads();
    }
    if (tid == 0) {
        x3[a] = sh[0][0];
        v1p[3 * a] = sh[1][0]; v1p[3 * a + 1] = sh[2][0]; v1p[3 * a + 2] = sh[3][0];
    }
}

__global__ void zero_counter_kernel(unsigned int *c) { *c = 0u; }

// flag |= 1 when any of the n doubles is NaN or infinite (one atomic per wave at most)
__global__ void nonfinite_kernel(const double *__restrict__ p, long long n, unsigned int *__restrict__ flag)
{
    int bad = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        bad |= !isfinite(p[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

// FP32 slabs from the FP64 ones in the FP32 sweep's fragment order: AX [panel][outer][K block
// of 16][lane][4], BX [panel][K block][lane][4]; lane l holds inner index l & 15, element j
// is at K = 16 kb + 4 j + (l >> 4).  BX = 1: the source is the BX slab [panel][K][16 gamma].
template <int BX>
__global__ void pack_f32v_kernel(const double *__restrict__ src, float *__restrict__ dst, long long npanel, int kp,
                                 int kpf)
{
    const int nkb = kpf / 16;
    const long long total = npanel * (BX ? 16LL : 256LL) * kpf;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(idx & 3), lane = (int)((idx >> 2) & 63);
        const long long rr = idx >> 8;     // (panel, outer, kb) or (panel, kb)
        const int kb = (int)(rr % nkb);
        const long long po = rr / nkb;
        const int K = kb * 16 + j * 4 + (lane >> 4), inner = lane & 15;
        double v = 0.0;
        if (K < kp) {
            if (BX) {
                v = src[po * (16LL * kp) + (long long)K * 16 + inner];
            } else {
                const int outer = (int)(po % 16);
                const long long panel = po / 16;
                v = src[panel * (256LL * kp) + (((long long)(outer >> 1) * (kp / 4) + K / 4) * 64 + (K & 3) * 16 + inner) * 2 +
                        (outer & 1)];
            }
        }
        dst[idx] = (float)v;
    }
}

// grid-stride launches: enough work groups to fill the GPU, at most 65536
static int grid_for(long long total)
{
    long long g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

template <int FK>
static void launch_pack_split(const double *axr, const double *axc, const double *bx, void *const *sl, int no, int nb,
                              int kp, int kps, const double *sc, hipStream_t s)
{
    constexpr int EL = SplitCfg<FK>::FMT == 0 ? 8 : 32;
    const long long ax_g = (long long)no * nb * nb * 256 * kps / EL, bx_g = (long long)no * no * nb * 16 * kps / EL;
    pack_split_kernel<FK, 0><<<grid_for(ax_g), 256, 0, s>>>(axr, (unsigned char *)sl[0], (int *)sl[3],
                                                            (long long)no * nb * nb, kp, kps, 1.0 / sc[0]);
    pack_split_kernel<FK, 0><<<grid_for(ax_g), 256, 0, s>>>(axc, (unsigned char *)sl[1], (int *)sl[4],
                                                            (long long)no * nb * nb, kp, kps, 1.0 / sc[0]);
    pack_split_kernel<FK, 1><<<grid_for(bx_g), 256, 0, s>>>(bx, (unsigned char *)sl[2], (int *)sl[5],
                                                            (long long)no * no * nb, kp, kps, 1.0 / sc[1]);
}

// ---------------------------------------------------------------------------------
// thin C ABI used by the C host layer
// ---------------------------------------------------------------------------------
extern "C" {

int ccx_hip_device_count(int *n) { return hipGetDeviceCount(n) == hipSuccess ? 0 : 1; }

int ccx_hip_malloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return 0;
    return hipMalloc(p, bytes) == hipSuccess ? 0 : 1;
}

int ccx_hip_free(void *p) { return (p == nullptr || hipFree(p) == hipSuccess) ? 0 : 1; }

int ccx_hip_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return 0;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream) != hipSuccess) return 1;
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : 1;
}

int ccx_hip_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return 0;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess) return 1;
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : 1;
}

int ccx_hip_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return 0;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess ? 0 : 1;
}

int ccx_hip_sync(void *stream) { return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : 1; }

int ccx_hip_cu_count(int *n)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    return hipDeviceGetAttribute(n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess ? 0 : 1;
}

const char *ccx_hip_last_error(void) { return hipGetErrorString(hipGetLastError()); }


// form 0: FP64 slabs (double); form 1: FP32 slabs (float).  voep is written only when non-null.
int ccx_hip_pack(int form, const double *vvvo, const double *t2, const double *vm, const double *voe, void *axr,
                 void *axc, void *bx, double *voep, int no, int nu, int nup, int kp, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const int nb = nup / TB;
    const long long nax = (long long)no * nb * nb * kp * 256;
    const int gbx = grid_for((long long)no * no * nb * kp * 16);
    if (form == 0) {
        pack_ax_kernel<0, double><<<grid_for(nax), 256, 0, s>>>(vvvo, t2, (double *)axr, no, nu, nb, kp, no);
        pack_ax_kernel<1, double><<<grid_for(nax), 256, 0, s>>>(vvvo, t2, (double *)axc, no, nu, nb, kp, no);
        pack_bx_kernel<double><<<gbx, 256, 0, s>>>(t2, vm, (double *)bx, no, nu, nb, kp);
    } else {
        pack_ax_kernel<0, float><<<grid_for(nax), 256, 0, s>>>(vvvo, t2, (float *)axr, no, nu, nb, kp, no);
        pack_ax_kernel<1, float><<<grid_for(nax), 256, 0, s>>>(vvvo, t2, (float *)axc, no, nu, nb, kp, no);
        pack_bx_kernel<float><<<gbx, 256, 0, s>>>(t2, vm, (float *)bx, no, nu, nb, kp);
    }
    if (voep) pack_voe_kernel<<<grid_for((long long)no * no * nup * nup), 256, 0, s>>>(voe, voep, no, nu, nup);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The FP64 panels of one occupied index (both AX orientations) from vvvo_x = VVVO(:,:,:,x) (nu^3) and
// t2_x = T2(:,:,:,x) (nu^2 no); axr_x / axc_x point at that index's panels.
int ccx_hip_pack_ax_one(const double *vvvo_x, const double *t2_x, double *axr_x, double *axc_x, int no, int nu,
                        int nup, int kp, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const int nb = nup / TB;
    const long long n = (long long)nb * nb * kp * 256;
    pack_ax_kernel<0, double><<<grid_for(n), 256, 0, s>>>(vvvo_x, t2_x, axr_x, no, nu, nb, kp, 1);
    pack_ax_kernel<1, double><<<grid_for(n), 256, 0, s>>>(vvvo_x, t2_x, axc_x, no, nu, nb, kp, 1);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The FP64 BX panels of one ordered pair (y,z) and, when voe_yz is non-null, its padded VOE block:
// t2_yz = T2(:,:,y,z), vm_zy = VM(:,:,z,y), voe_yz = VOE(:,:,y,z).
int ccx_hip_pack_pair(const double *t2_yz, const double *vm_zy, const double *voe_yz, double *bx_yz,
                      double *voep_yz, int no, int nu, int nup, int kp, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const int nb = nup / TB;
    pack_bx_pair_kernel<<<grid_for((long long)nb * kp * 16), 256, 0, s>>>(t2_yz, vm_zy, bx_yz, no, nu, nb, kp);
    if (voe_yz) pack_voe_kernel<<<grid_for((long long)nup * nup), 256, 0, s>>>(voe_yz, voep_yz, 1, nu, nup);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_ve_to_vvvo(const double *ve, double *vvvo_x, int nu, void *stream)
{
    ve_to_vvvo_kernel<<<grid_for((long long)nu * nu * nu), 256, 0, (hipStream_t)stream>>>(ve, vvvo_x, nu);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// form 0: FP64 slabs; 1: FP32 slabs in the FP32 sweep's order (K padded to kps, a multiple of
// 48); 2: split BF16 (split[] = axr_s, axc_s, bx_s, K padded to kps, W scaled by wscale)
// K padding of a form's slabs: FP32 48; split forms two K blocks of their MFMA (64 for bf16,
// 256 for fp8 e4m3); FP64 4
int ccx_hip_form_kpad(int form) { return form == 1 ? 48 : form == 3 ? 256 : form >= 2 ? 64 : 4; }

// form 0: FP64 slabs; 1: FP32 slabs in the FP32 sweep's order (K padded to kps, a multiple of
// 48); 2-4: split slices (split[] = axr_s, axc_s, bx_s and, for form 3, the exponent words
// axr_e, axc_e, bx_e; K padded to kps, W scaled by wscale)
int ccx_hip_triples(int form, const void *axr, const void *axc, const void *bx, const double *voep,
                    const double *eh, const double *ep, const int *tuples, const int *blocks, double *e_part,
                    double *v1_part, double *wout, unsigned int *counter, int no, int nu, int nup, int kp,
                    int ntup, int nblk, int grid, const void *const *split, int kps, double wscale, void *stream)
{
    static bool attr_set[5] = {false, false, false, false, false};
    const size_t lds_bytes = sizeof(double) * LDS_DOUBLES;
    if (form < 0 || form > 4) return 3;
    if (form >= 2 && (split == nullptr || kps % ccx_hip_form_kpad(form) != 0)) return 3;
    if (form == 1 && kps % 48 != 0) return 3;   // FP32 slabs: K padded to 48 (kps)
    const void *fns[5] = {(const void *)triples_kernel<double, 0>, (const void *)triples_kernel<float, 1>,
                          (const void *)triples_kernel<float, 2>, (const void *)triples_kernel<float, 3>,
                          (const void *)triples_kernel<float, 4>};
    if (!attr_set[form]) {
        if (hipFuncSetAttribute(fns[form], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
            return 2;
        attr_set[form] = true;
    }
    const bool sp = form >= 2;
    TriplesArgs g;
    g.axr = axr; g.axc = axc; g.bx = bx; g.voe = voep; g.eh = eh; g.ep = ep; g.tuples = tuples;
    g.blocks = blocks; g.e_part = e_part; g.v1_part = v1_part; g.wout = wout; g.counter = counter;
    g.no = no; g.nu = nu; g.nup = nup; g.nb = nup / TB; g.kp = kp; g.ntup = ntup; g.nblk = nblk;
    g.nitems = (long long)ntup * nblk;
    g.axr_s = sp ? split[0] : nullptr;
    g.axc_s = sp ? split[1] : nullptr;
    g.bx_s = sp ? split[2] : nullptr;
    g.axr_e = form == 3 ? (const int *)split[3] : nullptr;
    g.axc_e = form == 3 ? (const int *)split[4] : nullptr;
    g.bx_e = form == 3 ? (const int *)split[5] : nullptr;
    g.kps = kps;
    g.wscale = sp ? wscale : 1.0;
    {
        const char *env = getenv("CCX_ITEM_ORDER");
        g.block_major = (env && env[0] == 't') ? 0 : 1;
    }
    if (g.nitems == 0) return 0;
    if (grid <= 0) grid = 256;
    if (grid > g.nitems) grid = (int)g.nitems;
    hipStream_t s = (hipStream_t)stream;
    zero_counter_kernel<<<1, 1, 0, s>>>(counter);
    switch (form) {
    case 0: triples_kernel<double, 0><<<grid, NTHR, lds_bytes, s>>>(g); break;
    case 1: triples_kernel<float, 1><<<grid, NTHR, lds_bytes, s>>>(g); break;
    case 2: triples_kernel<float, 2><<<grid, NTHR, lds_bytes, s>>>(g); break;
    case 3: triples_kernel<float, 3><<<grid, NTHR, lds_bytes, s>>>(g); break;
    default: triples_kernel<float, 4><<<grid, NTHR, lds_bytes, s>>>(g); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Bytes per slab element of a split form's slices (BF16 x 3: 6, FP8 x 4: 4, BF16 x 4: 8); the FP8
// form adds one exponent word per 32 elements.
int ccx_hip_split_elem_bytes(int form) { return form == 2 ? 6 : form == 3 ? 4 : form == 4 ? 8 : 0; }

// Split slabs from the FP64 ones: power-of-two scales sigma_A >= max|AX|, sigma_B >= max|BX| (the
// fixed-point grid of the BF16 x 4 slabs; for every form they keep FP64 inputs of any range inside
// FP32's), then the slices of AX / sigma_A (both orientations) and BX / sigma_B; sl[] = axr_s, axc_s,
// bx_s, axr_e, axc_e, bx_e; *wscale = sigma_A sigma_B.
int ccx_hip_pack_split(int form, const double *axr, const double *axc, const double *bx, void *const *sl,
                       unsigned long long *dmax, int no, int nb, int kp, int kps, double *wscale, void *stream)
{
    if (form < 2 || form > 4 || kps % ccx_hip_form_kpad(form) != 0) return 3;
    hipStream_t s = (hipStream_t)stream;
    const long long n_ax = (long long)no * nb * nb * 256 * kp, n_bx = (long long)no * no * nb * 16 * kp;
    zero_counter_kernel<<<1, 1, 0, s>>>(reinterpret_cast<unsigned int *>(dmax));
    zero_counter_kernel<<<1, 1, 0, s>>>(reinterpret_cast<unsigned int *>(dmax) + 1);
    zero_counter_kernel<<<1, 1, 0, s>>>(reinterpret_cast<unsigned int *>(dmax) + 2);
    zero_counter_kernel<<<1, 1, 0, s>>>(reinterpret_cast<unsigned int *>(dmax) + 3);
    absmax_kernel<<<grid_for(n_ax), 256, 0, s>>>(axr, n_ax, dmax);
    absmax_kernel<<<grid_for(n_bx), 256, 0, s>>>(bx, n_bx, dmax + 1);
    unsigned long long h[2];
    if (hipMemcpyAsync(h, dmax, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return 1;
    double sc[2];
    for (int q = 0; q < 2; ++q) {
        double m;
        memcpy(&m, &h[q], sizeof(m));
        int e = 0;
        if (m > 0.0) frexp(m, &e);   // m < 2^e
        sc[q] = m > 0.0 ? ldexp(1.0, e) : 1.0;
    }
    if (form == 2) launch_pack_split<2>(axr, axc, bx, sl, no, nb, kp, kps, sc, s);
    else if (form == 3) launch_pack_split<3>(axr, axc, bx, sl, no, nb, kp, kps, sc, s);
    else launch_pack_split<4>(axr, axc, bx, sl, no, nb, kp, kps, sc, s);
    *wscale = sc[0] * sc[1];
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_finalize(const double *e_part, const double *v1_part, const int *tuples, const int *blocks,
                     const double *t1, double *v1t, double *v1, double *out, int no, int nu, int ntup, int nblk,
                     void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (ntup > 0) {
        reduce_v1_tuple_kernel<<<grid_for((long long)ntup * nu * 3), 256, 0, s>>>(v1_part, blocks, v1t, ntup,
                                                                                  nblk, nu);
        reduce_v1_kernel<<<(nu * no + 255) / 256, 256, 0, s>>>(v1t, tuples, v1, ntup, nu, no);
    } else {
        if (hipMemsetAsync(v1, 0, sizeof(double) * nu * no, s) != hipSuccess) return 1;
    }
    finalize_energy_kernel<<<1, 1024, 0, s>>>(e_part, (long long)ntup * nblk, t1, v1, no, nu, out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_t1wt3(const double *w, const double *voe6, double dijk, const double *ep, int nu, double *x3,
                  double *v1p, void *stream)
{
    if (nu < 1) return 0;
    t1wt3_kernel<<<nu, 256, 0, (hipStream_t)stream>>>(w, voe6, dijk, ep, nu, x3, v1p);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_lds_bytes(void) { return (int)(sizeof(double) * LDS_DOUBLES); }

int ccx_hip_pack_f32v(const double *axr, const double *axc, const double *bx, float *axr32, float *axc32,
                      float *bx32, int no, int nb, int kp, int kpf, void *stream)
{
    if (kpf % 48 != 0) return 3;
    hipStream_t s = (hipStream_t)stream;
    const long long nax = (long long)no * nb * nb * 256 * kpf, nbx = (long long)no * no * nb * 16 * kpf;
    pack_f32v_kernel<0><<<grid_for(nax), 256, 0, s>>>(axr, axr32, (long long)no * nb * nb, kp, kpf);
    pack_f32v_kernel<0><<<grid_for(nax), 256, 0, s>>>(axc, axc32, (long long)no * nb * nb, kp, kpf);
    pack_f32v_kernel<1><<<grid_for(nbx), 256, 0, s>>>(bx, bx32, (long long)no * no * nb, kp, kpf);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// *nonfinite = 1 when any element of the arrays p[0..narr) (n[q] doubles each) is NaN or infinite
int ccx_hip_nonfinite(const double *const *p, const size_t *n, int narr, unsigned int *d_flag, int *nonfinite,
                      void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    zero_counter_kernel<<<1, 1, 0, s>>>(d_flag);
    for (int q = 0; q < narr; ++q)
        if (n[q] > 0) nonfinite_kernel<<<grid_for((long long)n[q]), 256, 0, s>>>(p[q], (long long)n[q], d_flag);
    unsigned int h = 0;
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&h, d_flag, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return 1;
    *nonfinite = h != 0;
    return 0;
}

int ccx_hip_unpack_vvvo(const double *packed, double *full, int no, int nu, void *stream)
{
    unpack_vvvo_kernel<<<grid_for((long long)nu * nu * nu * no), 256, 0, (hipStream_t)stream>>>(packed, full, no,
                                                                                                 nu);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_event_create(void **ev)
{
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return 1;
    *ev = (void *)e;
    return 0;
}

int ccx_hip_event_destroy(void *ev) { return hipEventDestroy((hipEvent_t)ev) == hipSuccess ? 0 : 1; }

int ccx_hip_event_record(void *ev, void *stream)
{
    return hipEventRecord((hipEvent_t)ev, (hipStream_t)stream) == hipSuccess ? 0 : 1;
}

int ccx_hip_event_ms(void *start, void *stop, float *ms)
{
    if (hipEventSynchronize((hipEvent_t)stop) != hipSuccess) return 1;
    return hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop) == hipSuccess ? 0 : 1;
}

}  // extern "C"
