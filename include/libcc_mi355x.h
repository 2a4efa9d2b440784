/*
 * libcc_mi355x — closed-shell CCSD(T) triples correction on AMD MI355X (gfx950).
 *
 * Drop-in C ABI for the triples path of ryanolson/libcc.  Two layers:
 *
 *  1. Legacy Fortran-ABI entry points with the reference's names and argument
 *     lists: lowercase name + trailing underscore, every argument by reference,
 *     INTEGER = int64_t (the reference builds with 8-byte default integers,
 *     src/triples-ijk/Makefile:10), DOUBLE PRECISION = double, arrays are raw
 *     column-major host pointers.  A Fortran caller links against this library
 *     unchanged.  As in the reference these return nothing and keep the energies
 *     in process-global state; unlike the reference they never stop the process:
 *     a failure is reported by ccx_last_status().
 *
 *  2. Context-based entry points (ccx_*) returning an int status (CCX_OK = 0),
 *     for callers that can hold a handle, use device-resident inputs, or want the
 *     other arithmetic forms.
 *
 * Array layouts are the reference's (Fortran order, 1-based in Fortran terms):
 *   eh(no), ep(nu), t1(no,nu), t2(nu,nu,no,no), vm(no,nu,no,no), voe(nu,nu,no,no),
 *   vvvo(nu,nu,nu,no) unpacked, or the distributed layout (nutr, nu*no) with
 *   nutr = nu*(nu+1)/2 (lower-packed symmetric first pair; column c + nu*(i-1)).
 */
#ifndef LIBCC_MI355X_H
#define LIBCC_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define CCX_OK 0
#define CCX_ERR_ARG 1      /* bad dimension, index or null pointer           */
#define CCX_ERR_DEVICE 2   /* no usable GPU or a HIP call failed              */
#define CCX_ERR_ALLOC 3    /* device or host allocation failed                */
#define CCX_ERR_STATE 4    /* call out of order (no inputs loaded, no init)   */
#define CCX_ERR_INPUT 5    /* NaN/Inf in an input array                       */
#define CCX_ERR_FORM 6     /* arithmetic form not available                   */
#define CCX_ERR_IO 7       /* restart file missing, short or mis-framed       */
#define CCX_ERR_SCRATCH 8  /* problem larger than the static scratch array of ccx_energy_static */

/* arithmetic forms of the triples contraction; denominators, energies, v1 and all
 * reductions are FP64 in every form */
#define CCX_FORM_FP64 0    /* full FP64 (v_mfma_f64_16x16x4_f64)              */
#define CCX_FORM_FP32 1    /* FP32 operands, FP32 accumulation (v_mfma_f32_16x16x4_f32) */
/* Split precision (Ozaki-style): each FP64 operand element is carried as several low-precision
 * slices whose sum reproduces it; the slice products of levels i + j < slices run on the BF16 or
 * FP8 matrix cores with FP32 accumulation, W is formed in FP64, and the denominators, the epilogue
 * and every reduction stay FP64.
 *   BF16X3: 3 bf16 slices per element, each the bf16 rounding of what the previous ones left (24
 *           bits relative to the element itself), 6 products, one FP32 accumulator.
 *   FP8X4:  4 fp8 e4m3 digit slices per 32-element row group under the group's own power-of-two
 *           exponent (16 bits of the group maximum), 10 products on
 *           v_mfma_scale_f32_16x16x128_f8f6f4 with exponent and slice weight as block scales.
 *   BF16X4: 4 bf16 fixed-point digit slices under one power-of-two scale per slab (32 bits of the
 *           slab maximum), 10 products; level 0 in its own FP32 accumulator (its products share
 *           one grid and add exactly), the higher levels in a second one. */
#define CCX_FORM_BF16X3 2
#define CCX_FORM_FP8X4 3
#define CCX_FORM_BF16X4 4

const char *ccx_error_string(int status);
int ccx_last_status(void);          /* status of the last legacy (void) call    */

/* ---------------------------------------------------------------- context API */
typedef struct ccx_ctx ccx_ctx;

/* New context for no occupied / nu virtual orbitals on the current HIP device. */
int ccx_create(int64_t no, int64_t nu, ccx_ctx **ctx);
int ccx_destroy(ccx_ctx *ctx);

/* Upload inputs from host memory (reference layouts).  vvvo_layout: 0 = unpacked
 * (nu,nu,nu,no), 1 = distributed packed (nutr, nu*no).  Inputs are checked for
 * NaN/Inf (CCX_ERR_INPUT). */
int ccx_load_host(ccx_ctx *ctx, const double *eh, const double *ep, const double *t1,
                  const double *t2, const double *vm, const double *voe, const double *vvvo,
                  int vvvo_layout);

/* Same with device pointers (HBM-resident inputs; nothing crosses PCIe). */
int ccx_load_device(ccx_ctx *ctx, const double *eh, const double *ep, const double *t1,
                    const double *t2, const double *vm, const double *voe, const double *vvvo,
                    int vvvo_layout);

/* E[T] (etd), E_ST (ets) and v1(nu,no) over all i>=j>=k tuples in one persistent
 * launch.  v1 may be NULL; v1_on_device selects whether v1 is a device pointer. */
int ccx_energy(ccx_ctx *ctx, int form, double *ets, double *etd, double *v1, int v1_on_device);

/* ccx_energy over a subset of the tuples: n tuples (i,j,k), 1-based, i >= j >= k and not all
 * equal, three int64 each in ijk.  A tuple with i > j > k carries weight 1, (i,j,j) and (i,i,j)
 * weight 1/2, as in the full set (src/triples-ijk/t1wt3.f90:98-102, t3squa_acc.f:50-54), so the
 * energies and v1 are the partial sums the reference's driver accumulates over those tuples
 * (e.g. one tuple's ddcc_t_ijk_acc contribution, src/triples-ijk/ddcc_t_ijk_big.f:1-21).  At most
 * as many tuples as the full set holds. */
int ccx_energy_tuples(ccx_ctx *ctx, int form, int64_t n, const int64_t *ijk, double *ets, double *etd, double *v1,
                      int v1_on_device);

/* W_ijk(a,b,c) cube (nu^3, column-major, host) of one tuple, 1-based i,j,k, as the
 * reference's ijk_tuple leaves it in V3 (before the W(a,a,a)=0 step it is identical
 * off the diagonal). */
int ccx_tuple_w(ccx_ctx *ctx, int64_t i, int64_t j, int64_t k, double *w);

/* Virtual-triple view of W: v3(i,j,k) = W_ijk(a,b,c) for one virtual triple (1-based a,b,c)
 * over all occupied i,j,k (no^3, column-major, host), with v3(i,i,i) = 0.  This is the cube
 * the triples-abc prototype abc_tuple_formv3(a,b,c,...,v3) assembles from 12 GEMMs and
 * zerot3 (src/triples-abc/abc_tuple.f90:1-81).  That routine itself is not exported: it is
 * never called, its array arguments are declared with sizes its GEMMs do not use, and no
 * caller defines what its slices hold.  no^3 above 2^30 returns CCX_ERR_ARG (also for
 * ccx_energy_abc). */
int ccx_w_abc(ccx_ctx *ctx, int64_t a, int64_t b, int64_t c, double *v3);

/* E[T] (etd) re-summed over virtual triples: sum over a>b>c of X_abc plus half the sum of
 * X_abb and X_aab, X_abc = sum_ijk W [8W - 4(W_ikj + W_kji + W_jik) + 2(W_kij + W_jki)] / D
 * on the cubes of ccx_w_abc.  Equal to the etd of ccx_energy up to summation order; a
 * cross-check route (FP64 FMA, not the matrix cores), not the fast path. */
int ccx_energy_abc(ccx_ctx *ctx, double *etd);

/* Timing of the last ccx_energy call in milliseconds (HIP events around the
 * persistent kernel only, and around the whole call). */
int ccx_last_timing(const ccx_ctx *ctx, double *kernel_ms, double *total_ms);

/* Work items and the MFMA FLOPs the FP64 form's persistent launch executes (16-padded
 * virtual blocks, less the sweeps of waves whose batches are all padding), for roofline
 * accounting. */
int ccx_work_stats(const ccx_ctx *ctx, int64_t *items, double *executed_flops);

/* Adjoint of E(T) = etd + ets with respect to every input array (FP64), by re-running the
 * tuple loop: dE/dT1 (no,nu), dE/dT2 (nu,nu,no,no), dE/dVM (no,nu,no,no), dE/dVOE
 * (nu,nu,no,no) and dE/dVVVO (nu,nu,nu,no, unpacked; each entry an independent variable)
 * in the reference layouts; T2 and VOE are treated as independent inputs.  Any gradient
 * pointer may be NULL; on_device selects device or host pointers.  Also returns the
 * energies.  The reference has no adjoint; the energy it differentiates is
 * src/triples-ijk/triples.F90:29-313 (W from ijk_tuple.f90:56-84, E[T] and v1 from
 * t1wt3.f90:53-102, ets = 2 T1.v1 from triples.F90:310-313). */
int ccx_gradient(ccx_ctx *ctx, double *ets, double *etd, double *dt1, double *dt2, double *dvm, double *dvoe,
                 double *dvvvo, int on_device);

/* ccx_gradient with the arithmetic form of the per-tuple W recomputation: CCX_FORM_FP64 (as
 * ccx_gradient), CCX_FORM_FP32 (W from FP32 operands) or CCX_FORM_BF16X3 (W from the split
 * BF16 slices); both reduced forms accumulate W in FP32, and after it everything is FP64:
 * dE/dW, every back-contraction and every gradient accumulation.  Other forms return
 * CCX_ERR_FORM. */
int ccx_gradient_form(ccx_ctx *ctx, int form, double *ets, double *etd, double *dt1, double *dt2, double *dvm,
                      double *dvoe, double *dvvvo, int on_device);

/* Set the stream (hipStream_t cast to void*) later calls use; NULL = default stream. */
int ccx_set_stream(ccx_ctx *ctx, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the FP64 solve of
 * ccx_load_device + ccx_energy (unpacked VVVO) with every work buffer in one static device array
 * of CCX_STATIC_SCRATCH_BYTES per GPU (held until the process ends) instead of allocated memory.
 * All pointers are device pointers in the reference layouts; on return e[0] = etd, e[1] = ets and
 * v1(nu,no) are being written on `stream` (hipStream_t cast to void*, NULL = default stream) and
 * are ready when it is synchronized.  Results are bitwise those of ccx_energy.  Problems whose
 * buffers exceed the array return CCX_ERR_SCRATCH before any device work (no=4, nu=87 needs about
 * 58 MB; no=8, nu=120 does not fit).  NaN/Inf input returns CCX_ERR_INPUT as ccx_load_device does.
 * Every call on a GPU shares that GPU's array, so calls that overlap on one GPU (other host
 * threads, other streams) race and must be serialized by the caller. */
#define CCX_STATIC_SCRATCH_BYTES (64ull << 20)
int ccx_energy_static(int64_t no, int64_t nu, const double *eh, const double *ep, const double *t1,
                      const double *t2, const double *vm, const double *voe, const double *vvvo, double *e,
                      double *v1, void *stream);

/* ---------------------------------------------------------------- legacy Fortran ABI */

/* common_cc_init(nocc, nvir): src/common/common_cc.f90:342-369.  Sets the problem
 * dimensions and zeroes ets/etd for the calls below. */
void common_cc_init_(int64_t *nocc, int64_t *nvir);

/* DDI handle adapter: the GAMESS entry receives VVVO as a DDI array handle.  Register
 * the host array behind a handle first (layout as in ccx_load_host). */
void ccx_register_vvvo_(int64_t *handle, const double *vvvo, int64_t *layout);

/* cc_triples_acc(eh,ep,v1,t1,t2,v3,t3,vm,voe,vvvo_hnd,gms_ets,gms_etd):
 * src/triples-ijk/triples.F90:1-12,319-320.  Writes v1(nu,no), gms_ets, gms_etd and the
 * module energies; like the reference it transposes t1 in place to (nu,no) on return
 * (triples.F90:311).  v3 and t3 are scratch in the reference and are not touched. */
void cc_triples_acc_(double *eh, double *ep, double *v1, double *t1, double *t2, double *v3,
                     double *t3, double *vm, double *voe, int64_t *vvvo_hnd, double *gms_ets,
                     double *gms_etd);

/* cc_triples(eh,ep,v1,t1,t2,v3,t3,vm,voe,ve_i,ve_j,ve_k): src/triples-ijk/triples_cpu.F90:243-249
 * (and src/triples-abc/triples.f90:193-199).  VVVO comes from the handle registered for the
 * current d_vvvo (ccx_set_d_vvvo_); results in the module energies (ccx_get_ets_/etd_). */
void cc_triples_(double *eh, double *ep, double *v1, double *t1, double *t2, double *v3,
                 double *t3, double *vm, double *voe, double *ve_i, double *ve_j, double *ve_k);

/* The module variable d_vvvo (src/common/common_cc.f90:10) that cc_triples reads. */
void ccx_set_d_vvvo_(int64_t *handle);

/* Module accumulators ets / etd (src/common/common_cc.f90:32). */
double ccx_get_ets_(void);
double ccx_get_etd_(void);

/* ijk_tuple(i,j,k,t2_i,t2_j,t2_k,vm_ij,vm_ji,vm_ik,vm_ki,vm_jk,vm_kj,ve_i,ve_j,ve_k,v3):
 * src/triples-ijk/ijk_tuple.f90:34-35.  Forms W_ijk into v3 (nu^3) on the GPU; uses the
 * dimensions from common_cc_init_.  t2_x are T2(:,:,:,x) blocks, vm_xy VM(:,:,x,y) blocks
 * and ve_x the fetched VE_x(d,p,q) = VVVO(d,q,p,x) cubes, exactly as the reference passes them. */
void ijk_tuple_(int64_t *i, int64_t *j, int64_t *k, const double *t2_i, const double *t2_j,
                const double *t2_k, const double *vm_ij, const double *vm_ji, const double *vm_ik,
                const double *vm_ki, const double *vm_jk, const double *vm_kj, const double *ve_i,
                const double *ve_j, const double *ve_k, double *v3);

/* t1wt3_ijk(i,j,k,no,nu,v3,voe_ij,voe_ji,voe_ik,voe_ki,voe_jk,voe_kj,t1,eh,ep):
 * src/triples-ijk/t1wt3.f90:18-109.  From the tuple's W cube in v3 (nu^3), adds the tuple's E[T]
 * to the module etd (halved when i=j or j=k, t1wt3.f90:98-102) and its v1 rows i, j, k to t1,
 * which is the v1 accumulator (nu,no) as the driver passes it.  Runs on the GPU. */
void t1wt3_ijk_(int64_t *i, int64_t *j, int64_t *k, int64_t *no, int64_t *nu, const double *v3,
                const double *voe_ij, const double *voe_ji, const double *voe_ik, const double *voe_ki,
                const double *voe_jk, const double *voe_kj, double *t1, const double *eh, const double *ep);

/* ddcc_t_ijk_acc(no,nu,i,j,k,t1,t2,vm,v3,voe,eh,ep,ve_i,ve_j,ve_k): src/triples-ijk/ddcc_t_ijk_big.f:1-21.
 * One distinct tuple of the driver loop: W into v3, E[T] into the module etd, v1 rows into t1
 * (the v1 accumulator, nu x no).  Full t2, vm, voe; ve_x = VE_x(d,p,q) = VVVO(d,q,p,x). */
void ddcc_t_ijk_acc_(int64_t *no, int64_t *nu, int64_t *i, int64_t *j, int64_t *k, double *t1, double *t2,
                     double *vm, double *v3, double *voe, double *eh, double *ep, double *ve_i, double *ve_j,
                     double *ve_k);

/* ddcc_t_ijj_acc / ddcc_t_iij_acc(no,nu,i,j,t1,t2,vm,v3,t3,voe,o1,eh,ep,temp,ve_i,ve_j):
 * src/triples-ijk/ddcc_t_ijj_iij_big.f:18-94.  The degenerate tuples (i,j,j) and (i,i,j):
 * E[T] with weight 1/2 into etd, v1 rows into t1 (the accumulator), W into v3.  t3, o1
 * (the T1 amplitudes) and temp are the reference's scratch and are not read. */
void ddcc_t_ijj_acc_(int64_t *no, int64_t *nu, int64_t *i, int64_t *j, double *t1, double *t2, double *vm,
                     double *v3, double *t3, double *voe, double *o1, double *eh, double *ep, double *temp,
                     double *ve_i, double *ve_j);
void ddcc_t_iij_acc_(int64_t *no, int64_t *nu, int64_t *i, int64_t *j, double *t1, double *t2, double *vm,
                     double *v3, double *t3, double *voe, double *o1, double *eh, double *ep, double *temp,
                     double *ve_i, double *ve_j);

/* trant3 family: in-place permutations of an n^3 cube, new(x,y,z) = old(s(x,y,z)).
 *   trant3_1_ (x,z,y)  trant3_3_ (z,y,x)  trant3_4_ (y,z,x)  trant3_5_ (z,x,y)
 * src/common/common_cc.f90:62-156.  trant3_acc_(v,nu,id) and trant3_smp_(v,nu,id) select by
 * id 1..5: (x,z,y) (y,x,z) (z,y,x) (y,z,x) (z,x,y) (common_cc.f90:240-333,
 * src/triples-abc/trant3.f:1).  tranmd_23_ swaps indices 2 and 3 of an (n1,n2,n3,n4) array
 * (common_cc.f90:159-201).  The triples-abc copy also defines a one-argument trant3_1_
 * (src/triples-abc/ijk_tuple.f90:55); this library exports the two-argument form. */
void trant3_1_(int64_t *n, double *v);
void trant3_3_(int64_t *n, double *v);
void trant3_4_(int64_t *n, double *v);
void trant3_5_(int64_t *n, double *v);
void trant3_acc_(double *v, int64_t *nu, int64_t *id);
void trant3_smp_(double *v, int64_t *nu, int64_t *id);
void tranmd_23_(double *a, int64_t *n1, int64_t *n2, int64_t *n3, int64_t *n4);

/* trant3_1_async / trant3_4_async(acc_sync,n,v): src/triples-ijk/ijk_tuple_acc.f90:109-153, and
 * tranmd_23_acc(acc_sync,a,n1,n2,n3,n4): src/common/common_cc.f90:203-237.  The same
 * permutations as trant3_1_ / trant3_4_ / tranmd_23_ on host memory; the OpenACC queue id is
 * not used (the call completes before it returns). */
void trant3_1_async_(int64_t *acc_sync, int64_t *n, double *v);
void trant3_4_async_(int64_t *acc_sync, int64_t *n, double *v);
void tranmd_23_acc_(int64_t *acc_sync, double *a, int64_t *n1, int64_t *n2, int64_t *n3, int64_t *n4);

/* TRANMD_SMP(a,n1,n2,n3,n4,ij): src/triples-ijk/common_functions.f:141-421.  In-place
 * permutation of a(n1,n2,n3,n4), new(x1,x2,x3,x4) = old(source) by code ij:
 *   12 (x2,x1,x3,x4)  13 (x3,x2,x1,x4)  14 (x4,x2,x3,x1)  23 (x1,x3,x2,x4)  24 (x1,x4,x3,x2)
 *   34 (x1,x2,x4,x3)  231 (x2,x3,x1,x4)  312 (x3,x1,x2,x4)  341 (x3,x2,x4,x1)
 *   413 (x4,x2,x1,x3)  1234 (x4,x3,x2,x1)
 * The permuted extents must be equal (CCX_ERR_ARG otherwise, array unchanged); any other
 * code leaves the array as it is, like the reference. */
void tranmd_smp_(double *a, int64_t *n1, int64_t *n2, int64_t *n3, int64_t *n4, int64_t *ij);

/* DDCC_T_IJK_BIG(no,nu,i,j,k,t1,t2,vm,v3,t3,voe,o1,tmp,eh,ep,ve_i,ve_j,ve_k): the triples-abc
 * per-tuple routine, src/triples-abc/ddcc_t_ijk_big.f:1-32 (driver call
 * src/triples-abc/triples.f90:255 passes t1 = the v1 accumulator, o1 = T1).  Adds the distinct
 * tuple's E[T] to etd and its v1 rows to t1, the same sums as ddcc_t_ijk_acc_; v3 receives W;
 * t3, o1 and tmp are not read. */
void ddcc_t_ijk_big_(int64_t *no, int64_t *nu, int64_t *i, int64_t *j, int64_t *k, double *t1, double *t2,
                     double *vm, double *v3, double *t3, double *voe, double *o1, double *tmp, double *eh, double *ep,
                     double *ve_i, double *ve_j, double *ve_k);

/* The reference's CUDA tuple path: ijk_gpu_init(no,nu,eh,ep,v1,t2,vm,voe),
 * ijk_gpu_driver(nu,no,i,j,k,ve_i,ve_j,ve_k,t1,eh,ep,etd), ijk_gpu_finalize(no,nu,etd)
 * (call sites src/triples-ijk/triples_cpu.F90:363,497 and ddcc_t_ijk_big_cuda.f:12-14).  init
 * keeps the caller's arrays and clears the accumulators; driver takes each occupied index's
 * VVVO patch as the distributed array holds it (nutr x nu: column c is the lower-packed (p,q)
 * block of VVVO(:,:,c,x)) and adds the distinct tuple's E[T] and v1 rows to the accumulators;
 * finalize adds them to *etd and to the v1 given to init.  ccx_last_status() reports errors. */
void ijk_gpu_init_(int64_t *no, int64_t *nu, double *eh, double *ep, double *v1, double *t2, double *vm,
                   double *voe);
void ijk_gpu_driver_(int64_t *nu, int64_t *no, int64_t *i, int64_t *j, int64_t *k, double *ve_i, double *ve_j,
                     double *ve_k, double *t1, double *eh, double *ep, double *etd);
void ijk_gpu_finalize_(int64_t *no, int64_t *nu, double *etd);

/* ---------------------------------------------------------------- triples.restart */

/* The reference's input file (src/triples-ijk/program.F90:44-46 reads no and nu,
 * cc_triples_readinp :195-218 the arrays): Fortran unformatted sequential records, each
 * framed by a 4-byte byte count before and after, integers int64.  Records: no; nu; eh(no);
 * ep(nu); t1(no*nu); t2(no^2 nu^2); vm(no^3 nu); voe(no^2 nu^2); then no*nu records of
 * nutr = nu(nu+1)/2 doubles, record j = column j = c + nu*x of the lower-packed VVVO (rows
 * p(p+1)/2 + q, q <= p, 0-based).  All arrays in the reference layouts.  Returns CCX_ERR_IO
 * for a missing, short or mis-framed file. */
int ccx_restart_dims(const char *path, int64_t *no, int64_t *nu);
/* Reads every array; vvvo_packed receives (nutr, no*nu). */
int ccx_read_restart(const char *path, int64_t no, int64_t nu, double *eh, double *ep, double *t1, double *t2,
                     double *vm, double *voe, double *vvvo_packed);
/* Writes the file; vvvo in either layout of ccx_load_host (0 = (nu,nu,nu,no), 1 = packed). */
int ccx_write_restart(const char *path, int64_t no, int64_t nu, const double *eh, const double *ep,
                      const double *t1, const double *t2, const double *vm, const double *voe, const double *vvvo,
                      int vvvo_layout);
/* ccx_load_host from a restart file whose no, nu match the context's. */
int ccx_load_restart(ccx_ctx *ctx, const char *path);

#ifdef __cplusplus
}
#endif

#endif /* LIBCC_MI355X_H */
