/*
 * Host layer (C) of libcc_mi355x: argument checking, the device-resident data model,
 * the tuple / block schedule, and the legacy Fortran-ABI entry points of the
 * reference's triples path.  All device work goes through the thin ABI in ccx_dev.h.
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/libcc_mi355x.h"
#include "ccx_dev.h"

#define TB 16

struct ccx_ctx {
    int64_t no, nu, nup, nb, kp;
    int64_t ntup, nblk;
    void *stream;
    int loaded;
    double *d_axr, *d_axc, *d_bx, *d_voep, *d_eh, *d_ep, *d_t1;
    float *d_axr32, *d_axc32, *d_bx32; /* FP32 form slabs, converted on first use after a load */
    int64_t kpf;                        /* K padded to 48 for the FP32 form */
    int f32_ready;
    void *d_axr_s, *d_axc_s, *d_bx_s;  /* split forms: slices of each element, built on first use */
    int *d_axr_e, *d_axc_e, *d_bx_e;   /* FP8 form: per-lane block exponent words */
    size_t split_ax_bytes, split_bx_bytes;   /* allocated sizes of the slice slabs */
    unsigned long long *d_max;
    int64_t kps;                        /* K padding of the split form whose slices are built */
    double split_scale;
    int split_ready, split_form;
    int *d_tuples, *d_blocks;
    unsigned int *d_counter;
    unsigned int *d_flag;               /* NaN/Inf flag of ccx_load_device */
    double *d_epart, *d_v1part, *d_v1t, *d_v1, *d_out;
    int *h_tuples, *h_blocks;
    double kernel_ms, total_ms;
    void *ev[4];
};

const char *ccx_error_string(int s)
{
    switch (s) {
    case CCX_OK: return "ok";
    case CCX_ERR_ARG: return "invalid argument";
    case CCX_ERR_DEVICE: return "GPU unavailable or HIP call failed";
    case CCX_ERR_ALLOC: return "allocation failed";
    case CCX_ERR_STATE: return "call out of order";
    case CCX_ERR_INPUT: return "non-finite value in an input array";
    case CCX_ERR_FORM: return "arithmetic form not available";
    case CCX_ERR_IO: return "file missing, unreadable or not in the restart layout";
    case CCX_ERR_SCRATCH: return "problem larger than the static scratch array";
    default: return "unknown status";
    }
}

/* ------------------------------------------------------------------ schedule tables */

/* Occupied tuples in the reference driver's order: distinct i>j>k with i outermost
 * (DDCC_T_TASK, src/triples-ijk/common_functions.f:62-86), then for each i>j the
 * (i,j,j) and (i,i,j) tuples (src/triples-ijk/triples.F90:237-300).  0-based; the 4th
 * entry is 1 for a distinct tuple, 2 for (i,j,j) and 3 for (i,i,j) (degenerate tuples
 * carry energy weight 1/2, src/triples-ijk/t1wt3.f90:188-192). */
static int64_t build_tuples(int64_t no, int *out)
{
    int64_t n = 0;
    for (int64_t i = 0; i < no; ++i)
        for (int64_t j = 0; j < i; ++j)
            for (int64_t k = 0; k < j; ++k) {
                if (out) { out[4 * n] = (int)i; out[4 * n + 1] = (int)j; out[4 * n + 2] = (int)k; out[4 * n + 3] = 1; }
                ++n;
            }
    for (int64_t i = 0; i < no; ++i)
        for (int64_t j = 0; j < i; ++j) {
            if (out) {
                out[4 * n] = (int)i; out[4 * n + 1] = (int)j; out[4 * n + 2] = (int)j; out[4 * n + 3] = 2;
                out[4 * n + 4] = (int)i; out[4 * n + 5] = (int)i; out[4 * n + 6] = (int)j; out[4 * n + 7] = 3;
            }
            n += 2;
        }
    return n;
}

/* Virtual 16^3 block orbits A<=B<=C; 4th entry = how many positional permutations
 * map the block onto itself (1, 2 or 6), the factor by which the orbit covers the cube. */
static int64_t build_blocks(int64_t nb, int *out)
{
    int64_t n = 0;
    for (int64_t a = 0; a < nb; ++a)
        for (int64_t b = a; b < nb; ++b)
            for (int64_t c = b; c < nb; ++c) {
                if (out) {
                    int m = (a == b && b == c) ? 6 : (a == b || b == c) ? 2 : 1;
                    out[4 * n] = (int)a; out[4 * n + 1] = (int)b; out[4 * n + 2] = (int)c; out[4 * n + 3] = m;
                }
                ++n;
            }
    return n;
}

/* ------------------------------------------------------------------ context */

static void free_dev(ccx_ctx *c)
{
    double **d[] = {&c->d_axr, &c->d_axc, &c->d_bx, &c->d_voep, &c->d_eh, &c->d_ep, &c->d_t1,
                    &c->d_epart, &c->d_v1part, &c->d_v1t, &c->d_v1, &c->d_out};
    for (size_t q = 0; q < sizeof(d) / sizeof(d[0]); ++q) {
        ccx_hip_free(*d[q]);
        *d[q] = NULL;
    }
    ccx_hip_free(c->d_tuples);
    ccx_hip_free(c->d_blocks);
    ccx_hip_free(c->d_counter);
    ccx_hip_free(c->d_axr32);
    ccx_hip_free(c->d_axc32);
    ccx_hip_free(c->d_bx32);
    ccx_hip_free(c->d_axr_s);
    ccx_hip_free(c->d_axc_s);
    ccx_hip_free(c->d_bx_s);
    ccx_hip_free(c->d_max);
    ccx_hip_free(c->d_axr_e);
    ccx_hip_free(c->d_axc_e);
    ccx_hip_free(c->d_bx_e);
    c->d_axr_e = c->d_axc_e = c->d_bx_e = NULL;
    c->split_ax_bytes = c->split_bx_bytes = 0;
    c->d_tuples = c->d_blocks = NULL;
    c->d_counter = NULL;
    c->d_axr32 = c->d_axc32 = c->d_bx32 = NULL;
    c->d_axr_s = c->d_axc_s = c->d_bx_s = NULL;
    c->d_max = NULL;
    c->f32_ready = 0;
    c->split_ready = 0;
    c->loaded = 0;
}

int ccx_create(int64_t no, int64_t nu, ccx_ctx **out)
{
    if (!out) return CCX_ERR_ARG;
    *out = NULL;
    if (no < 1 || nu < 1 || no > 2048 || nu > 4096) return CCX_ERR_ARG;
    int ndev = 0;
    if (ccx_hip_device_count(&ndev) != 0 || ndev < 1) return CCX_ERR_DEVICE;
    ccx_ctx *c = (ccx_ctx *)calloc(1, sizeof(ccx_ctx));
    if (!c) return CCX_ERR_ALLOC;
    c->no = no;
    c->nu = nu;
    c->nup = (nu + TB - 1) / TB * TB;
    c->nb = c->nup / TB;
    c->kp = (nu + no + 3) / 4 * 4;
    c->kpf = (nu + no + 47) / 48 * 48;
    c->ntup = build_tuples(no, NULL);
    c->nblk = build_blocks(c->nb, NULL);
    c->h_tuples = (int *)malloc(sizeof(int) * 4 * (size_t)(c->ntup > 0 ? c->ntup : 1));
    c->h_blocks = (int *)malloc(sizeof(int) * 4 * (size_t)c->nblk);
    if (!c->h_tuples || !c->h_blocks) { free(c->h_tuples); free(c->h_blocks); free(c); return CCX_ERR_ALLOC; }
    build_tuples(no, c->h_tuples);
    build_blocks(c->nb, c->h_blocks);
    for (int q = 0; q < 4; ++q)
        if (ccx_hip_event_create(&c->ev[q]) != 0) { free(c->h_tuples); free(c->h_blocks); free(c); return CCX_ERR_DEVICE; }
    *out = c;
    return CCX_OK;
}

int ccx_destroy(ccx_ctx *c)
{
    if (!c) return CCX_ERR_ARG;
    free_dev(c);
    ccx_hip_free(c->d_flag);
    for (int q = 0; q < 4; ++q) ccx_hip_event_destroy(c->ev[q]);
    free(c->h_tuples);
    free(c->h_blocks);
    free(c);
    return CCX_OK;
}

int ccx_set_stream(ccx_ctx *c, void *stream)
{
    if (!c) return CCX_ERR_ARG;
    c->stream = stream;
    return CCX_OK;
}

static int finite_all(const double *p, size_t n)
{
    for (size_t q = 0; q < n; ++q)
        if (!isfinite(p[q])) return 0;
    return 1;
}

#define DALLOC(ptr, count)                                                             \
    do {                                                                               \
        if (ccx_hip_malloc((void **)&(ptr), sizeof(*(ptr)) * (size_t)(count)) != 0) { \
            rc = CCX_ERR_ALLOC;                                                        \
            goto fail;                                                                 \
        }                                                                              \
    } while (0)

/* Build the packed operand slabs and small arrays from device-resident raw inputs. */
static int build_from_device(ccx_ctx *c, const double *d_eh, const double *d_ep_raw, const double *d_t1,
                             const double *d_t2, const double *d_vm, const double *d_voe,
                             const double *d_vvvo_full)
{
    int rc = CCX_OK;
    const int64_t no = c->no, nu = c->nu, nup = c->nup, nb = c->nb, kp = c->kp;
    const int64_t nitems = c->ntup * c->nblk;
    double *h_ep = NULL;
    if (c->d_axr) goto allocated; /* buffers depend only on the dimensions: reuse */
    DALLOC(c->d_axr, no * nb * nb * kp * 256);
    DALLOC(c->d_axc, no * nb * nb * kp * 256);
    DALLOC(c->d_counter, 8);
    DALLOC(c->d_bx, no * no * kp * nup);
    DALLOC(c->d_voep, no * no * nup * nup);
    DALLOC(c->d_eh, no);
    DALLOC(c->d_ep, nup);
    DALLOC(c->d_t1, no * nu);
    DALLOC(c->d_tuples, 4 * (c->ntup > 0 ? c->ntup : 1));
    DALLOC(c->d_blocks, 4 * c->nblk);
    DALLOC(c->d_epart, nitems > 0 ? nitems : 1);
    DALLOC(c->d_v1part, (nitems > 0 ? nitems : 1) * 144);
    DALLOC(c->d_v1t, (c->ntup > 0 ? c->ntup : 1) * nu * 3);
    DALLOC(c->d_v1, no * nu);
    DALLOC(c->d_out, 4);
allocated:
    c->loaded = 0;
    c->f32_ready = 0;
    c->split_ready = 0;
    if (ccx_hip_pack(0, d_vvvo_full, d_t2, d_vm, d_voe, c->d_axr, c->d_axc, c->d_bx, c->d_voep, (int)no, (int)nu,
                     (int)nup, (int)kp, c->stream) != 0) { rc = CCX_ERR_DEVICE; goto fail; }
    /* eps_virtual padded to nup with finite values (padded elements carry W = 0) */
    h_ep = (double *)malloc(sizeof(double) * (size_t)nup);
    if (!h_ep) { rc = CCX_ERR_ALLOC; goto fail; }
    if (ccx_hip_memcpy_d2h(h_ep, d_ep_raw, sizeof(double) * (size_t)nu, c->stream) != 0) { rc = CCX_ERR_DEVICE; goto fail; }
    for (int64_t a = nu; a < nup; ++a) h_ep[a] = 1.0e3;
    if (ccx_hip_memcpy_h2d(c->d_ep, h_ep, sizeof(double) * (size_t)nup, c->stream) != 0 ||
        ccx_hip_memcpy_d2d(c->d_eh, d_eh, sizeof(double) * (size_t)no, c->stream) != 0 ||
        ccx_hip_memcpy_d2d(c->d_t1, d_t1, sizeof(double) * (size_t)(no * nu), c->stream) != 0 ||
        ccx_hip_memcpy_h2d(c->d_tuples, c->h_tuples, sizeof(int) * 4 * (size_t)c->ntup, c->stream) != 0 ||
        ccx_hip_memcpy_h2d(c->d_blocks, c->h_blocks, sizeof(int) * 4 * (size_t)c->nblk, c->stream) != 0 ||
        ccx_hip_sync(c->stream) != 0) { rc = CCX_ERR_DEVICE; goto fail; }
    free(h_ep);
    c->loaded = 1;
    return CCX_OK;
fail:
    free(h_ep);
    free_dev(c);
    return rc;
}

int ccx_load_device(ccx_ctx *c, const double *eh, const double *ep, const double *t1, const double *t2,
                    const double *vm, const double *voe, const double *vvvo, int vvvo_layout)
{
    if (!c || !eh || !ep || !t1 || !t2 || !vm || !voe || !vvvo) return CCX_ERR_ARG;
    if (vvvo_layout != 0 && vvvo_layout != 1) return CCX_ERR_ARG;
    {   /* the same contract as ccx_load_host: NaN/Inf anywhere is CCX_ERR_INPUT */
        const int64_t no = c->no, nu = c->nu;
        const double *arr[7] = {eh, ep, t1, t2, vm, voe, vvvo};
        const size_t n[7] = {(size_t)no, (size_t)nu, (size_t)(no * nu), (size_t)(nu * nu * no * no),
                             (size_t)(no * nu * no * no), (size_t)(nu * nu * no * no),
                             vvvo_layout == 0 ? (size_t)(nu * nu * nu * no) : (size_t)(nu * (nu + 1) / 2 * nu * no)};
        int bad = 0;
        if (!c->d_flag && ccx_hip_malloc((void **)&c->d_flag, sizeof(unsigned int)) != 0) return CCX_ERR_ALLOC;
        if (ccx_hip_nonfinite(arr, n, 7, c->d_flag, &bad, c->stream) != 0) return CCX_ERR_DEVICE;
        if (bad) return CCX_ERR_INPUT;
    }
    if (vvvo_layout == 0) return build_from_device(c, eh, ep, t1, t2, vm, voe, vvvo);
    double *full = NULL;
    if (ccx_hip_malloc((void **)&full, sizeof(double) * (size_t)(c->nu * c->nu * c->nu * c->no)) != 0)
        return CCX_ERR_ALLOC;
    int rc = ccx_hip_unpack_vvvo(vvvo, full, (int)c->no, (int)c->nu, c->stream) != 0 ? CCX_ERR_DEVICE : CCX_OK;
    if (rc == CCX_OK) rc = build_from_device(c, eh, ep, t1, t2, vm, voe, full);
    ccx_hip_sync(c->stream);
    ccx_hip_free(full);
    return rc;
}

int ccx_load_host(ccx_ctx *c, const double *eh, const double *ep, const double *t1, const double *t2,
                  const double *vm, const double *voe, const double *vvvo, int vvvo_layout)
{
    if (!c || !eh || !ep || !t1 || !t2 || !vm || !voe || !vvvo) return CCX_ERR_ARG;
    if (vvvo_layout != 0 && vvvo_layout != 1) return CCX_ERR_ARG;
    const int64_t no = c->no, nu = c->nu;
    const size_t n_t2 = (size_t)(nu * nu * no * no), n_vm = (size_t)(no * nu * no * no);
    const size_t n_vvvo = vvvo_layout == 0 ? (size_t)(nu * nu * nu * no) : (size_t)(nu * (nu + 1) / 2 * nu * no);
    if (!finite_all(eh, (size_t)no) || !finite_all(ep, (size_t)nu) || !finite_all(t1, (size_t)(no * nu)) ||
        !finite_all(t2, n_t2) || !finite_all(vm, n_vm) || !finite_all(voe, n_t2) || !finite_all(vvvo, n_vvvo))
        return CCX_ERR_INPUT;
    int rc = CCX_OK;
    double *d_eh = NULL, *d_ep = NULL, *d_t1 = NULL, *d_t2 = NULL, *d_vm = NULL, *d_voe = NULL, *d_vvvo = NULL;
    DALLOC(d_eh, no);
    DALLOC(d_ep, nu);
    DALLOC(d_t1, no * nu);
    DALLOC(d_t2, n_t2);
    DALLOC(d_vm, n_vm);
    DALLOC(d_voe, n_t2);
    DALLOC(d_vvvo, n_vvvo);
    if (ccx_hip_memcpy_h2d(d_eh, eh, sizeof(double) * (size_t)no, c->stream) != 0 ||
        ccx_hip_memcpy_h2d(d_ep, ep, sizeof(double) * (size_t)nu, c->stream) != 0 ||
        ccx_hip_memcpy_h2d(d_t1, t1, sizeof(double) * (size_t)(no * nu), c->stream) != 0 ||
        ccx_hip_memcpy_h2d(d_t2, t2, sizeof(double) * n_t2, c->stream) != 0 ||
        ccx_hip_memcpy_h2d(d_vm, vm, sizeof(double) * n_vm, c->stream) != 0 ||
        ccx_hip_memcpy_h2d(d_voe, voe, sizeof(double) * n_t2, c->stream) != 0 ||
        ccx_hip_memcpy_h2d(d_vvvo, vvvo, sizeof(double) * n_vvvo, c->stream) != 0) {
        rc = CCX_ERR_DEVICE;
        goto fail;
    }
    rc = ccx_load_device(c, d_eh, d_ep, d_t1, d_t2, d_vm, d_voe, d_vvvo, vvvo_layout);
    ccx_hip_sync(c->stream);
fail:
    ccx_hip_free(d_eh); ccx_hip_free(d_ep); ccx_hip_free(d_t1); ccx_hip_free(d_t2);
    ccx_hip_free(d_vm); ccx_hip_free(d_voe); ccx_hip_free(d_vvvo);
    return rc;
}

static int grid_size(const ccx_ctx *c, int64_t nitems)
{
    int cus = 256;
    if (ccx_hip_cu_count(&cus) != 0 || cus < 1) cus = 256;
    int64_t g = cus;
    if (g > nitems) g = nitems;
    return (int)(g < 1 ? 1 : g);
    (void)c;
}

/* FP32 slabs from the FP64 ones (same layout), allocated on first use */
static int ensure_f32(ccx_ctx *c)
{
    if (c->f32_ready) return CCX_OK;
    const int64_t nax = c->no * c->nb * c->nb * 256 * c->kpf, nbx = c->no * c->no * c->nb * 16 * c->kpf;
    if (!c->d_axr32) {
        if (ccx_hip_malloc((void **)&c->d_axr32, sizeof(float) * (size_t)nax) != 0 ||
            ccx_hip_malloc((void **)&c->d_axc32, sizeof(float) * (size_t)nax) != 0 ||
            ccx_hip_malloc((void **)&c->d_bx32, sizeof(float) * (size_t)nbx) != 0) {
            ccx_hip_free(c->d_axr32); ccx_hip_free(c->d_axc32); ccx_hip_free(c->d_bx32);
            c->d_axr32 = c->d_axc32 = c->d_bx32 = NULL;
            return CCX_ERR_ALLOC;
        }
    }
    if (ccx_hip_pack_f32v(c->d_axr, c->d_axc, c->d_bx, c->d_axr32, c->d_axc32, c->d_bx32, (int)c->no, (int)c->nb,
                          (int)c->kp, (int)c->kpf, c->stream) != 0)
        return CCX_ERR_DEVICE;
    c->f32_ready = 1;
    return CCX_OK;
}

/* the slices of a split form (CCX_FORM_BF16X3, _FP8X4, _BF16X4) from the FP64 slabs, built on the
 * first call in that form after a load; the slice slabs are reallocated when a form needs more
 * room than the last one (the FP8 form's exponent words are allocated with its first use) */
static int ensure_split(ccx_ctx *c, int form)
{
    if (c->split_ready && c->split_form == form) return CCX_OK;
    const int64_t pad = ccx_hip_form_kpad(form), eb = ccx_hip_split_elem_bytes(form);
    const int64_t kps = (c->nu + c->no + pad - 1) / pad * pad;
    const size_t ax_bytes = (size_t)(c->no * c->nb * c->nb * 256 * kps * eb);
    const size_t bx_bytes = (size_t)(c->no * c->no * c->nb * 16 * kps * eb);
    c->split_ready = 0;
    if (!c->d_axr_s || ax_bytes > c->split_ax_bytes || bx_bytes > c->split_bx_bytes) {
        ccx_hip_free(c->d_axr_s); ccx_hip_free(c->d_axc_s); ccx_hip_free(c->d_bx_s); ccx_hip_free(c->d_max);
        c->d_axr_s = c->d_axc_s = c->d_bx_s = NULL;
        c->d_max = NULL;
        c->split_ax_bytes = c->split_bx_bytes = 0;
        if (ccx_hip_malloc(&c->d_axr_s, ax_bytes) != 0 || ccx_hip_malloc(&c->d_axc_s, ax_bytes) != 0 ||
            ccx_hip_malloc(&c->d_bx_s, bx_bytes) != 0 ||
            ccx_hip_malloc((void **)&c->d_max, 2 * sizeof(unsigned long long)) != 0) {
            ccx_hip_free(c->d_axr_s); ccx_hip_free(c->d_axc_s); ccx_hip_free(c->d_bx_s); ccx_hip_free(c->d_max);
            c->d_axr_s = c->d_axc_s = c->d_bx_s = NULL;
            c->d_max = NULL;
            return CCX_ERR_ALLOC;
        }
        c->split_ax_bytes = ax_bytes;
        c->split_bx_bytes = bx_bytes;
    }
    if (form == CCX_FORM_FP8X4 && !c->d_axr_e) {   /* one word per 32 elements */
        const size_t ax_w = (size_t)(c->no * c->nb * c->nb * 256 * kps / 32), bx_w = (size_t)(c->no * c->no * c->nb * 16 * kps / 32);
        if (ccx_hip_malloc((void **)&c->d_axr_e, sizeof(int) * ax_w) != 0 ||
            ccx_hip_malloc((void **)&c->d_axc_e, sizeof(int) * ax_w) != 0 ||
            ccx_hip_malloc((void **)&c->d_bx_e, sizeof(int) * bx_w) != 0) {
            ccx_hip_free(c->d_axr_e); ccx_hip_free(c->d_axc_e); ccx_hip_free(c->d_bx_e);
            c->d_axr_e = c->d_axc_e = c->d_bx_e = NULL;
            return CCX_ERR_ALLOC;
        }
    }
    void *sl[6] = {c->d_axr_s, c->d_axc_s, c->d_bx_s, c->d_axr_e, c->d_axc_e, c->d_bx_e};
    if (ccx_hip_pack_split(form, c->d_axr, c->d_axc, c->d_bx, sl, c->d_max, (int)c->no, (int)c->nb, (int)c->kp,
                           (int)kps, &c->split_scale, c->stream) != 0)
        return CCX_ERR_DEVICE;
    c->kps = kps;
    c->split_form = form;
    c->split_ready = 1;
    return CCX_OK;
}

static int is_split_form(int form) { return form == CCX_FORM_BF16X3 || form == CCX_FORM_FP8X4 || form == CCX_FORM_BF16X4; }

/* One persistent launch over the tuple table d_tup (ntup entries of 4 ints) and the
 * deterministic reductions; energies to ets and etd, v1 (nu,no) to v1 when non-null. */
static int energy_run(ccx_ctx *c, int form, const int *d_tup, int64_t ntup, double *ets, double *etd, double *v1,
                      int v1_on_device)
{
    const int64_t nitems = ntup * c->nblk;
    double host_out[2] = {0.0, 0.0};
    if (form == CCX_FORM_FP32) {
        int rc = ensure_f32(c);
        if (rc != CCX_OK) return rc;
    }
    if (is_split_form(form)) {
        int rc = ensure_split(c, form);
        if (rc != CCX_OK) return rc;
    }
    const void *split[6] = {c->d_axr_s, c->d_axc_s, c->d_bx_s, c->d_axr_e, c->d_axc_e, c->d_bx_e};
    const void *axr = form == CCX_FORM_FP32 ? (const void *)c->d_axr32 : (const void *)c->d_axr;
    const void *axc = form == CCX_FORM_FP32 ? (const void *)c->d_axc32 : (const void *)c->d_axc;
    const void *bx = form == CCX_FORM_FP32 ? (const void *)c->d_bx32 : (const void *)c->d_bx;
    if (ccx_hip_event_record(c->ev[0], c->stream) != 0) return CCX_ERR_DEVICE;
    if (ccx_hip_event_record(c->ev[1], c->stream) != 0) return CCX_ERR_DEVICE;
    if (nitems > 0 &&
        ccx_hip_triples(form, axr, axc, bx, c->d_voep, c->d_eh, c->d_ep, d_tup, c->d_blocks, c->d_epart,
                        c->d_v1part, NULL, c->d_counter, (int)c->no, (int)c->nu, (int)c->nup, (int)c->kp,
                        (int)ntup, (int)c->nblk, grid_size(c, nitems), split,
                        form == CCX_FORM_FP32 ? (int)c->kpf : (int)c->kps, c->split_scale, c->stream) != 0)
        return CCX_ERR_DEVICE;
    if (ccx_hip_event_record(c->ev[2], c->stream) != 0) return CCX_ERR_DEVICE;
    if (ccx_hip_finalize(c->d_epart, c->d_v1part, d_tup, c->d_blocks, c->d_t1, c->d_v1t, c->d_v1, c->d_out,
                         (int)c->no, (int)c->nu, (int)ntup, (int)c->nblk, c->stream) != 0)
        return CCX_ERR_DEVICE;
    if (ccx_hip_event_record(c->ev[3], c->stream) != 0) return CCX_ERR_DEVICE;
    if (ccx_hip_memcpy_d2h(host_out, c->d_out, sizeof(double) * 2, c->stream) != 0) return CCX_ERR_DEVICE;
    if (v1) {
        const size_t bytes = sizeof(double) * (size_t)(c->no * c->nu);
        int r = v1_on_device ? ccx_hip_memcpy_d2d(v1, c->d_v1, bytes, c->stream)
                             : ccx_hip_memcpy_d2h(v1, c->d_v1, bytes, c->stream);
        if (r != 0 || ccx_hip_sync(c->stream) != 0) return CCX_ERR_DEVICE;
    }
    float k_ms = 0.0f, t_ms = 0.0f;
    if (ccx_hip_event_ms(c->ev[1], c->ev[2], &k_ms) != 0 || ccx_hip_event_ms(c->ev[0], c->ev[3], &t_ms) != 0)
        return CCX_ERR_DEVICE;
    c->kernel_ms = k_ms;
    c->total_ms = t_ms;
    if (etd) *etd = host_out[0];
    if (ets) *ets = host_out[1];
    return CCX_OK;
}

static int form_known(int form)
{
    return form == CCX_FORM_FP64 || form == CCX_FORM_FP32 || is_split_form(form);
}

int ccx_energy(ccx_ctx *c, int form, double *ets, double *etd, double *v1, int v1_on_device)
{
    if (!c) return CCX_ERR_ARG;
    if (!c->loaded) return CCX_ERR_STATE;
    if (!form_known(form)) return CCX_ERR_FORM;
    return energy_run(c, form, c->d_tuples, c->ntup, ets, etd, v1, v1_on_device);
}

int ccx_energy_tuples(ccx_ctx *c, int form, int64_t n, const int64_t *ijk, double *ets, double *etd, double *v1,
                      int v1_on_device)
{
    if (!c || (n > 0 && !ijk) || n < 0) return CCX_ERR_ARG;
    if (!c->loaded) return CCX_ERR_STATE;
    if (!form_known(form)) return CCX_ERR_FORM;
    if (n > c->ntup) return CCX_ERR_ARG;   /* the partial buffers hold one full tuple set */
    int *h = (int *)malloc(sizeof(int) * 4 * (size_t)(n > 0 ? n : 1));
    if (!h) return CCX_ERR_ALLOC;
    for (int64_t t = 0; t < n; ++t) {
        const int64_t i = ijk[3 * t], j = ijk[3 * t + 1], k = ijk[3 * t + 2];
        if (i < 1 || k < 1 || i > c->no || !(i >= j && j >= k) || (i == j && j == k)) { free(h); return CCX_ERR_ARG; }
        h[4 * t] = (int)(i - 1); h[4 * t + 1] = (int)(j - 1); h[4 * t + 2] = (int)(k - 1);
        h[4 * t + 3] = (i > j && j > k) ? 1 : (j == k ? 2 : 3);
    }
    int *d = NULL;
    int rc = CCX_OK;
    if (ccx_hip_malloc((void **)&d, sizeof(int) * 4 * (size_t)(n > 0 ? n : 1)) != 0) { free(h); return CCX_ERR_ALLOC; }
    if (n > 0 && ccx_hip_memcpy_h2d(d, h, sizeof(int) * 4 * (size_t)n, c->stream) != 0) rc = CCX_ERR_DEVICE;
    if (rc == CCX_OK) rc = energy_run(c, form, d, n, ets, etd, v1, v1_on_device);
    ccx_hip_sync(c->stream);
    ccx_hip_free(d);
    free(h);
    return rc;
}

int ccx_tuple_w(ccx_ctx *c, int64_t i, int64_t j, int64_t k, double *w)
{
    if (!c || !w) return CCX_ERR_ARG;
    if (!c->loaded) return CCX_ERR_STATE;
    if (i < 1 || j < 1 || k < 1 || i > c->no || j > c->no || k > c->no) return CCX_ERR_ARG;
    const size_t nu3 = (size_t)(c->nu * c->nu * c->nu);
    int rc = CCX_OK;
    int *d_tup = NULL;
    double *d_w = NULL, *d_ep1 = NULL, *d_vp = NULL;
    int tup[4] = {(int)(i - 1), (int)(j - 1), (int)(k - 1), 1};
    DALLOC(d_tup, 4);
    DALLOC(d_w, nu3);
    DALLOC(d_ep1, c->nblk);
    DALLOC(d_vp, c->nblk * 144);
    if (ccx_hip_memcpy_h2d(d_tup, tup, sizeof(tup), c->stream) != 0) { rc = CCX_ERR_DEVICE; goto fail; }
    if (ccx_hip_triples(0, c->d_axr, c->d_axc, c->d_bx, c->d_voep, c->d_eh, c->d_ep, d_tup, c->d_blocks, d_ep1,
                        d_vp, d_w, c->d_counter, (int)c->no, (int)c->nu, (int)c->nup, (int)c->kp, 1, (int)c->nblk,
                        grid_size(c, c->nblk), NULL, 0, 1.0, c->stream) != 0 ||
        ccx_hip_memcpy_d2h(w, d_w, sizeof(double) * nu3, c->stream) != 0)
        rc = CCX_ERR_DEVICE;
fail:
    ccx_hip_free(d_tup); ccx_hip_free(d_w); ccx_hip_free(d_ep1); ccx_hip_free(d_vp);
    return rc;
}

/* ------------------------------------------------------------------ virtual triples */

/* The virtual-triple kernels index an occupied cube with 32-bit ints. */
#define ABC_MAX_CUBE ((int64_t)1 << 30)

/* The virtual triples a >= b >= c, not all equal, in the order a > b > c (a outermost), then
 * (a,b,b) and (a,a,b) for each a > b; kind 1, 2, 3 as for occupied tuples.  Writes entries
 * [first, first + n) of that list into out (4 ints each) and returns the list length. */
static int64_t abc_triples(int64_t nu, int64_t first, int64_t n, int *out)
{
    int64_t pos = 0;
#define ABC_PUT(A, B, C, KIND)                                                                 \
    do {                                                                                      \
        if (pos >= first && pos < first + n) {                                                \
            int *q = out + 4 * (pos - first);                                                 \
            q[0] = (int)(A); q[1] = (int)(B); q[2] = (int)(C); q[3] = (KIND);                 \
        }                                                                                     \
        ++pos;                                                                                \
    } while (0)
    for (int64_t a = 0; a < nu; ++a)
        for (int64_t b = 0; b < a; ++b) {
            if (!out || pos + b <= first) { pos += b; continue; }   /* count or skip a whole c row */
            for (int64_t c = 0; c < b; ++c) ABC_PUT(a, b, c, 1);
        }
    for (int64_t a = 0; a < nu; ++a)
        for (int64_t b = 0; b < a; ++b) {
            ABC_PUT(a, b, b, 2);
            ABC_PUT(a, a, b, 3);
        }
#undef ABC_PUT
    return pos;
}

int ccx_w_abc(ccx_ctx *c, int64_t a, int64_t b, int64_t cc, double *v3)
{
    if (!c || !v3) return CCX_ERR_ARG;
    if (!c->loaded) return CCX_ERR_STATE;
    if (a < 1 || b < 1 || cc < 1 || a > c->nu || b > c->nu || cc > c->nu) return CCX_ERR_ARG;
    if (c->no * c->no * c->no > ABC_MAX_CUBE) return CCX_ERR_ARG;
    const size_t no3 = (size_t)(c->no * c->no * c->no);
    int rc = CCX_OK;
    int *d_trip = NULL;
    double *d_w = NULL;
    int trip[4] = {(int)(a - 1), (int)(b - 1), (int)(cc - 1), 1};
    DALLOC(d_trip, 4);
    DALLOC(d_w, no3);
    if (ccx_hip_memcpy_h2d(d_trip, trip, sizeof(trip), c->stream) != 0 ||
        ccx_hip_w_abc(c->d_axr, c->d_bx, d_trip, 1, (int)c->no, (int)c->nu, (int)c->nb, (int)c->kp, d_w,
                      c->stream) != 0 ||
        ccx_hip_memcpy_d2h(v3, d_w, sizeof(double) * no3, c->stream) != 0)
        rc = CCX_ERR_DEVICE;
fail:
    ccx_hip_free(d_trip); ccx_hip_free(d_w);
    return rc;
}

int ccx_energy_abc(ccx_ctx *c, double *etd)
{
    if (!c || !etd) return CCX_ERR_ARG;
    if (!c->loaded) return CCX_ERR_STATE;
    const int64_t no3 = c->no * c->no * c->no;
    if (no3 > ABC_MAX_CUBE) return CCX_ERR_ARG;
    const int64_t ntrip = abc_triples(c->nu, 0, 0, NULL);
    int64_t batch = ((int64_t)256 << 20) / (no3 * (int64_t)sizeof(double));
    if (batch < 1) batch = 1;
    if (batch > 65536) batch = 65536;
    if (batch > ntrip) batch = ntrip;
    int rc = CCX_OK;
    int *h_trip = NULL, *d_trip = NULL;
    double *d_w = NULL, *d_x = NULL, *h_x = NULL;
    double sum = 0.0;
    if (ntrip > 0) {
        h_trip = (int *)malloc(sizeof(int) * 4 * (size_t)batch);
        h_x = (double *)malloc(sizeof(double) * (size_t)batch);
        if (!h_trip || !h_x) { rc = CCX_ERR_ALLOC; goto fail; }
        DALLOC(d_trip, 4 * batch);
        DALLOC(d_w, no3 * batch);
        DALLOC(d_x, batch);
    }
    /* batches in list order; the partials of each batch are added on the host in order */
    for (int64_t first = 0; first < ntrip; first += batch) {
        const int64_t n = ntrip - first < batch ? ntrip - first : batch;
        abc_triples(c->nu, first, n, h_trip);
        if (ccx_hip_memcpy_h2d(d_trip, h_trip, sizeof(int) * 4 * (size_t)n, c->stream) != 0 ||
            ccx_hip_w_abc(c->d_axr, c->d_bx, d_trip, (int)n, (int)c->no, (int)c->nu, (int)c->nb, (int)c->kp, d_w,
                          c->stream) != 0 ||
            ccx_hip_x_abc(d_w, d_trip, (int)n, (int)c->no, c->d_eh, c->d_ep, d_x, c->stream) != 0 ||
            ccx_hip_memcpy_d2h(h_x, d_x, sizeof(double) * (size_t)n, c->stream) != 0) {
            rc = CCX_ERR_DEVICE;
            goto fail;
        }
        for (int64_t t = 0; t < n; ++t) sum += h_x[t];
    }
    *etd = sum;
fail:
    free(h_trip); free(h_x);
    ccx_hip_free(d_trip); ccx_hip_free(d_w); ccx_hip_free(d_x);
    return rc;
}

int ccx_last_timing(const ccx_ctx *c, double *kernel_ms, double *total_ms)
{
    if (!c) return CCX_ERR_ARG;
    if (kernel_ms) *kernel_ms = c->kernel_ms;
    if (total_ms) *total_ms = c->total_ms;
    return CCX_OK;
}

int ccx_work_stats(const ccx_ctx *c, int64_t *items, double *executed_flops)
{
    if (!c) return CCX_ERR_ARG;
    const int64_t nitems = c->ntup * c->nblk;
    if (items) *items = nitems;
    /* per item: 6 MFMA terms x 16^3 outputs x kp MACs per accumulated ordering; orderings:
     * 6 (distinct tuple, A<B<C), 3 (two equal blocks, or a degenerate tuple on distinct
     * blocks), 2 (a degenerate tuple on two equal blocks), 1 (A=B=C).
     * FP64 form: of the 8 waves (2 batches each) a phase runs only those whose batches reach
     * a virtual index of the phase's batch block (C in phase 0, B in phases 1 and 2, with a
     * ragged C of three distinct blocks moved to role B). */
    double units = 0.0;
    for (int64_t t = 0; t < c->ntup; ++t)
        for (int64_t q = 0; q < c->nblk; ++q) {
            const int kind = c->h_tuples[4 * t + 3], mult = c->h_blocks[4 * q + 3];
            int bb = c->h_blocks[4 * q + 1], bc = c->h_blocks[4 * q + 2];
            if (mult == 1 && c->nu - 16 * bc < 16) { const int x = bb; bb = bc; bc = x; }
            double fr[2];
            const int blk[2] = {bc, bb};
            for (int r = 0; r < 2; ++r) {
                int real = (int)(c->nu - 16 * blk[r]);
                if (real > 16) real = 16;
                fr[r] = (double)((real + 1) / 2) / 8.0;
            }
            const double ord = mult == 6 ? 1.0 : kind != 1 ? (mult == 1 ? 3.0 : 2.0) : mult == 1 ? 6.0 : 3.0;
            units += ord * (fr[0] + 2.0 * fr[1]) / 3.0;
        }
    if (executed_flops) *executed_flops = units * 6.0 * 4096.0 * 2.0 * (double)c->kp;
    return CCX_OK;
}

/* ------------------------------------------------------------------ adjoint */

/* sigma-terms of W (x position, AX index axes u,v, BX pair positions y,z, BX axis wf):
 * W_t(e) = sum_sigma sum_K AX[t[xp]](e_u,e_v,K) BX[t[yp],t[zp]](e_wf,K); same table as
 * the device code */
static const int SIGMA_H[6][6] = {{0, 0, 1, 1, 2, 2}, {1, 1, 0, 0, 2, 2}, {2, 2, 1, 1, 0, 0},
                                  {0, 0, 2, 2, 1, 1}, {1, 1, 2, 2, 0, 0}, {2, 2, 0, 0, 1, 1}};
static const int PERM_H[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
/* the sigma-term each term becomes when two tuple positions are swapped: 1,2 for (i,j,j),
 * 0,1 for (i,i,j) (each SIGMA_H row with those positions exchanged) */
static const int SIGMA_PAIR_IJJ[6] = {3, 5, 4, 0, 2, 1};
static const int SIGMA_PAIR_IIJ[6] = {1, 0, 5, 4, 3, 2};

/* CSR lists of (tl, code) entries grouped by key, keys in increasing order, entries in
 * (tl, code) order: list[u], off[u..u+1] for the nonzero keys only.  Returns #keys. */
/* panel splits of the gbx back-contraction: enough work groups to fill the GPU when a
 * batch holds few entries */
static int gbx_nsplit(int64_t nent, int64_t nb)
{
    int64_t ns = nent > 0 ? (2048 + nent - 1) / nent : 1;
    if (ns > nb) ns = nb;
    return ns < 1 ? 1 : (int)ns;
}

static int build_csr(int nkeys, int n, const int *key, const int *tl, const int *code, int *list, int *off,
                     int *ent, int *count)
{
    memset(count, 0, sizeof(int) * (size_t)nkeys);
    for (int q = 0; q < n; ++q) count[key[q]]++;
    int nu_ = 0, pos = 0;
    for (int k = 0; k < nkeys; ++k)
        if (count[k]) {
            list[nu_] = k;
            off[nu_] = pos;
            pos += count[k];
            count[k] = off[nu_];    /* running insert position */
            ++nu_;
        }
    off[nu_] = pos;
    for (int q = 0; q < n; ++q) {   /* stable: entries keep (tl, code) order */
        const int at = count[key[q]]++;
        ent[2 * at] = tl[q];
        ent[2 * at + 1] = code[q];
    }
    return nu_;
}

int ccx_gradient(ccx_ctx *c, double *ets, double *etd, double *dt1, double *dt2, double *dvm, double *dvoe,
                 double *dvvvo, int on_device)
{
    return ccx_gradient_form(c, CCX_FORM_FP64, ets, etd, dt1, dt2, dvm, dvoe, dvvvo, on_device);
}

int ccx_gradient_form(ccx_ctx *c, int form, double *ets, double *etd, double *dt1, double *dt2, double *dvm,
                      double *dvoe, double *dvvvo, int on_device)
{
    if (!c) return CCX_ERR_ARG;
    if (!c->loaded) return CCX_ERR_STATE;
    if (form != CCX_FORM_FP64 && form != CCX_FORM_FP32 && form != CCX_FORM_BF16X3) return CCX_ERR_FORM;
    if (form == CCX_FORM_BF16X3) {
        int rs = ensure_split(c, form);
        if (rs != CCX_OK) return rs;
    }
    if (form == CCX_FORM_FP32) {
        int rs = ensure_f32(c);
        if (rs != CCX_OK) return rs;
    }
    const void *w32[3] = {form == CCX_FORM_FP32 ? (const void *)c->d_axr32 : c->d_axr_s,
                          form == CCX_FORM_FP32 ? (const void *)c->d_axc32 : c->d_axc_s,
                          form == CCX_FORM_FP32 ? (const void *)c->d_bx32 : c->d_bx_s};
    const int64_t no = c->no, nu = c->nu, nup = c->nup, nb = c->nb, kp = c->kp;
    const int64_t ntup = c->ntup, nblk = c->nblk;
    int rc = CCX_OK;
    double *d_gax = NULL, *d_gbx = NULL, *d_dvoep = NULL, *d_wbar = NULL, *d_slots = NULL, *d_wstage = NULL;
    double *d_gpart = NULL;
    int64_t gpart_units = 0;   /* entries x splits the gbx partial buffer holds */
    double *d_out[5] = {NULL, NULL, NULL, NULL, NULL};
    int *d_lists = NULL, *h_lists = NULL, *h_tmp = NULL;
    double host_out[2] = {0.0, 0.0};
    /* batch of tuples: Wbar cubes + dVOE slots within a memory budget (CCX_ADJ_BATCH_MB) */
    const double per_tuple = 8.0 * (double)(nup * nup * nup) + 8.0 * 6.0 * (double)(nb * nb * nb) * 256.0;
    double budget = 16384.0;
    const char *env = getenv("CCX_ADJ_BATCH_MB");
    if (env && atof(env) > 0) budget = atof(env);
    int64_t tb = (int64_t)(budget * 1048576.0 / per_tuple);
    if (tb < 1) tb = 1;
    if (tb > ntup) tb = ntup > 0 ? ntup : 1;
    const int64_t nent = 6 * tb;                   /* entries per list kind per batch */
    const int64_t nkey = no * no;
    /* host list scratch: key/tl/code arrays (3 x nent), count (nkey), and three CSR sets */
    const int64_t lsz = 3 * (nkey + (nkey + 1) + 2 * nent);
    h_tmp = (int *)malloc(sizeof(int) * (size_t)(3 * nent + nkey));
    h_lists = (int *)malloc(sizeof(int) * (size_t)lsz);
    if (!h_tmp || !h_lists) { rc = CCX_ERR_ALLOC; goto fail; }
    DALLOC(d_gax, no * nup * nup * kp);
    DALLOC(d_gbx, no * no * nup * kp);
    DALLOC(d_dvoep, no * no * nup * nup);
    DALLOC(d_wbar, tb * nup * nup * nup);
    DALLOC(d_slots, tb * 6 * nb * nb * nb * 256);
    DALLOC(d_lists, lsz);
    DALLOC(d_wstage, (int64_t)grid_size(c, tb * nblk) * (int64_t)ccx_hip_adj_stage_doubles());
    {   /* gbx partial tiles: the largest entries x splits any batch can need (a batch holds
         * 1 .. 6 tb gBX entries, fewer than 6 per tuple when degenerate tuples merge pairs,
         * and the split count grows as the entry count falls) */
        int64_t umax = 1;
        for (int64_t n = 1; n <= 6 * tb; ++n) {
            const int64_t u = n * gbx_nsplit(n, nb);
            if (u > umax) umax = u;
        }
        gpart_units = umax;
        DALLOC(d_gpart, (int64_t)ccx_hip_gbx_part_doubles((int)nup, (int)kp, (int)umax, 1));
    }
    if (ccx_hip_zero(d_gax, sizeof(double) * (size_t)(no * nup * nup * kp), c->stream) != 0 ||
        ccx_hip_zero(d_gbx, sizeof(double) * (size_t)(no * no * nup * kp), c->stream) != 0 ||
        ccx_hip_zero(d_dvoep, sizeof(double) * (size_t)(no * no * nup * nup), c->stream) != 0) {
        rc = CCX_ERR_DEVICE;
        goto fail;
    }
    if (ccx_hip_event_record(c->ev[0], c->stream) != 0) { rc = CCX_ERR_DEVICE; goto fail; }
    for (int64_t t0 = 0; t0 < ntup; t0 += tb) {
        const int tcount = (int)(t0 + tb <= ntup ? tb : ntup - t0);
        int *key = h_tmp, *tlv = h_tmp + nent, *code = h_tmp + 2 * nent, *cnt = h_tmp + 3 * nent;
        /* three CSR sets laid out back to back: [list nkey][off nkey+1][ent 2 nent] */
        const int64_t set = nkey + (nkey + 1) + 2 * nent;
        int nkeys[3], nents[3];
        for (int kind = 0; kind < 3; ++kind) {
            int n = 0;
            for (int tl = 0; tl < tcount; ++tl) {
                const int *tp = c->h_tuples + 4 * (t0 + tl);
                /* A degenerate tuple's Wbar is symmetric in the two virtual axes of its equal
                 * positions, so the sigma-terms pair up (positions 1,2 swapped for (i,j,j), 0,1
                 * for (i,i,j)) into identical back-contractions: keep one of each pair with
                 * weight 2 (code bit 3) in the gAX / gBX lists. */
                const int *pair = tp[3] == 2 ? SIGMA_PAIR_IJJ : tp[3] == 3 ? SIGMA_PAIR_IIJ : NULL;
                for (int q = 0; q < 6; ++q) {
                    int cq = q;
                    if (kind < 2 && pair) {
                        if (pair[q] < q) continue;
                        cq = q | 8;
                    }
                    if (kind == 0) key[n] = tp[SIGMA_H[q][0]];                               /* gAX: x */
                    else if (kind == 1) key[n] = tp[SIGMA_H[q][3]] * (int)no + tp[SIGMA_H[q][4]]; /* gBX: (y,z) */
                    else key[n] = tp[PERM_H[q][1]] * (int)no + tp[PERM_H[q][2]];           /* dVOE: (q,r) */
                    tlv[n] = tl;
                    code[n] = cq;
                    ++n;
                }
            }
            int *base = h_lists + kind * set;
            nkeys[kind] = build_csr((int)nkey, n, key, tlv, code, base, base + nkey, base + nkey + (nkey + 1), cnt);
            nents[kind] = n;
            /* the kernels index entries 0 .. n-1 only: the CSR total must be exactly n */
            if (base[nkey + nkeys[kind]] != n) { rc = CCX_ERR_STATE; goto fail; }
        }
        if ((int64_t)nents[1] * gbx_nsplit(nents[1], nb) > gpart_units) { rc = CCX_ERR_STATE; goto fail; }
        if (ccx_hip_memcpy_h2d(d_lists, h_lists, sizeof(int) * (size_t)lsz, c->stream) != 0) { rc = CCX_ERR_DEVICE; goto fail; }
        const int *dl[3][3];
        for (int kind = 0; kind < 3; ++kind) {
            dl[kind][0] = d_lists + kind * set;
            dl[kind][1] = dl[kind][0] + nkey;
            dl[kind][2] = dl[kind][1] + (nkey + 1);
        }
        if (ccx_hip_adj_items(c->d_axr, c->d_axc, c->d_bx, c->d_voep, c->d_eh, c->d_ep, c->d_t1, c->d_tuples,
                              c->d_blocks, c->d_epart, c->d_v1part, d_wbar, d_slots, d_wstage, c->d_counter, (int)no,
                              (int)nu,
                              (int)nup, (int)kp, (int)ntup, (int)nblk, (int)t0, tcount,
                              grid_size(c, (int64_t)tcount * nblk), form, form == CCX_FORM_FP64 ? NULL : w32,
                              form == CCX_FORM_FP32 ? (int)c->kpf : (int)c->kps, c->split_scale, c->stream) != 0 ||
            ccx_hip_gax(d_wbar, c->d_bx, d_gax, c->d_tuples, (int)t0, dl[0][0], dl[0][1], dl[0][2], nkeys[0], (int)no,
                        (int)nup, (int)kp, c->stream) != 0 ||
            ccx_hip_gbx(d_wbar, c->d_axr, c->d_axc, d_gbx, d_gpart, c->d_tuples, (int)t0, dl[1][0], dl[1][1],
                        dl[1][2], nkeys[1], nents[1], gbx_nsplit(nents[1], nb), (int)no, (int)nup, (int)kp,
                        c->stream) != 0 ||
            ccx_hip_dvoe_reduce(d_slots, d_dvoep, dl[2][0], dl[2][1], dl[2][2], nkeys[2], (int)no, (int)nup,
                                c->stream) != 0) {
            rc = CCX_ERR_DEVICE;
            goto fail;
        }
    }
    if (ccx_hip_finalize(c->d_epart, c->d_v1part, c->d_tuples, c->d_blocks, c->d_t1, c->d_v1t, c->d_v1, c->d_out,
                         (int)no, (int)nu, (int)ntup, (int)nblk, c->stream) != 0) { rc = CCX_ERR_DEVICE; goto fail; }
    {
        double *outs[5] = {dvvvo, dt2, dvm, dvoe, dt1};
        const int64_t sizes[5] = {nu * nu * nu * no, nu * nu * no * no, no * nu * no * no, nu * nu * no * no, no * nu};
        double *dst[5];
        for (int q = 0; q < 5; ++q) {
            dst[q] = NULL;
            if (!outs[q]) continue;
            if (on_device) dst[q] = outs[q];
            else { DALLOC(d_out[q], sizes[q]); dst[q] = d_out[q]; }
        }
        if (ccx_hip_unpack_grad(d_gax, d_gbx, d_dvoep, c->d_v1, dst[0], dst[1], dst[2], dst[3], dst[4], (int)no,
                                (int)nu, (int)nup, (int)kp, c->stream) != 0 ||
            ccx_hip_event_record(c->ev[3], c->stream) != 0 ||
            ccx_hip_memcpy_d2h(host_out, c->d_out, sizeof(double) * 2, c->stream) != 0) {
            rc = CCX_ERR_DEVICE;
            goto fail;
        }
        for (int q = 0; q < 5; ++q)
            if (outs[q] && !on_device &&
                ccx_hip_memcpy_d2h(outs[q], d_out[q], sizeof(double) * (size_t)sizes[q], c->stream) != 0) {
                rc = CCX_ERR_DEVICE;
                goto fail;
            }
    }
    {
        float t_ms = 0.0f;
        if (ccx_hip_event_ms(c->ev[0], c->ev[3], &t_ms) != 0) { rc = CCX_ERR_DEVICE; goto fail; }
        c->kernel_ms = t_ms;
        c->total_ms = t_ms;
    }
    if (etd) *etd = host_out[0];
    if (ets) *ets = host_out[1];
fail:
    ccx_hip_sync(c->stream);
    ccx_hip_free(d_gax); ccx_hip_free(d_gbx); ccx_hip_free(d_dvoep); ccx_hip_free(d_wbar); ccx_hip_free(d_slots);
    ccx_hip_free(d_lists);
    ccx_hip_free(d_wstage);
    ccx_hip_free(d_gpart);
    for (int q = 0; q < 5; ++q) ccx_hip_free(d_out[q]);
    free(h_tmp);
    free(h_lists);
    return rc;
}

/* ------------------------------------------------------------------ legacy Fortran ABI */

static int64_t g_no = 0, g_nu = 0;
static double g_ets = 0.0, g_etd = 0.0;
static int g_status = CCX_OK;
static int64_t g_d_vvvo = 0;
static ccx_ctx *g_ctx = NULL;

#define NREG 16
static struct { int64_t handle; const double *ptr; int64_t layout; } g_reg[NREG];
static int g_nreg = 0;

int ccx_last_status(void) { return g_status; }

void common_cc_init_(int64_t *nocc, int64_t *nvir)
{
    if (!nocc || !nvir || *nocc < 1 || *nvir < 1) { g_status = CCX_ERR_ARG; return; }
    g_no = *nocc;
    g_nu = *nvir;
    g_ets = 0.0;
    g_etd = 0.0;
    g_status = CCX_OK;
}

void ccx_register_vvvo_(int64_t *handle, const double *vvvo, int64_t *layout)
{
    if (!handle || !vvvo || !layout || (*layout != 0 && *layout != 1)) { g_status = CCX_ERR_ARG; return; }
    for (int q = 0; q < g_nreg; ++q)
        if (g_reg[q].handle == *handle) { g_reg[q].ptr = vvvo; g_reg[q].layout = *layout; g_status = CCX_OK; return; }
    if (g_nreg == NREG) { g_status = CCX_ERR_ALLOC; return; }
    g_reg[g_nreg].handle = *handle;
    g_reg[g_nreg].ptr = vvvo;
    g_reg[g_nreg].layout = *layout;
    ++g_nreg;
    g_status = CCX_OK;
}

void ccx_set_d_vvvo_(int64_t *handle)
{
    if (!handle) { g_status = CCX_ERR_ARG; return; }
    g_d_vvvo = *handle;
    g_status = CCX_OK;
}

double ccx_get_ets_(void) { return g_ets; }
double ccx_get_etd_(void) { return g_etd; }

static int legacy_ctx(void)
{
    if (g_no < 1 || g_nu < 1) return CCX_ERR_STATE;
    if (g_ctx && (g_ctx->no != g_no || g_ctx->nu != g_nu)) {
        ccx_destroy(g_ctx);
        g_ctx = NULL;
    }
    if (!g_ctx) return ccx_create(g_no, g_nu, &g_ctx);
    return CCX_OK;
}

/* TRPOSE(t1,tmp,no,nu,1): t1(no,nu) -> t1(nu,no) in place (src/triples-ijk/common_functions.f:3-41),
 * through the caller's scratch b of n*m doubles */
static void transpose_inplace(double *a, double *b, int64_t n, int64_t m)
{
    for (int64_t j = 0; j < m; ++j)
        for (int64_t i = 0; i < n; ++i) b[j + m * i] = a[i + n * j];
    memcpy(a, b, sizeof(double) * (size_t)(n * m));
}

static void legacy_triples(int64_t handle, double *eh, double *ep, double *v1, double *t1, double *t2,
                           double *vm, double *voe, double *gms_ets, double *gms_etd)
{
    const double *vvvo = NULL;
    int64_t layout = 0;
    for (int q = 0; q < g_nreg; ++q)
        if (g_reg[q].handle == handle) { vvvo = g_reg[q].ptr; layout = g_reg[q].layout; }
    if (!vvvo) { g_status = CCX_ERR_STATE; return; }
    if (!eh || !ep || !v1 || !t1 || !t2 || !vm || !voe) { g_status = CCX_ERR_ARG; return; }
    /* scratch for the closing t1 transpose, taken first: if it is not there nothing is done */
    double *tmp = (double *)malloc(sizeof(double) * (size_t)(g_no > 0 && g_nu > 0 ? g_no * g_nu : 1));
    if (!tmp) { g_status = CCX_ERR_ALLOC; return; }
    int rc = legacy_ctx();
    if (rc == CCX_OK) rc = ccx_load_host(g_ctx, eh, ep, t1, t2, vm, voe, vvvo, (int)layout);
    double ets = 0.0, etd = 0.0;
    if (rc == CCX_OK) rc = ccx_energy(g_ctx, CCX_FORM_FP64, &ets, &etd, v1, 0);
    g_status = rc;
    if (rc == CCX_OK) {
        g_ets = ets;
        g_etd = etd;
        if (gms_ets) *gms_ets = ets;
        if (gms_etd) *gms_etd = etd;
        transpose_inplace(t1, tmp, g_no, g_nu);
    }
    free(tmp);
}

void cc_triples_acc_(double *eh, double *ep, double *v1, double *t1, double *t2, double *v3, double *t3,
                     double *vm, double *voe, int64_t *vvvo_hnd, double *gms_ets, double *gms_etd)
{
    (void)v3;
    (void)t3;
    if (!vvvo_hnd) { g_status = CCX_ERR_ARG; return; }
    legacy_triples(*vvvo_hnd, eh, ep, v1, t1, t2, vm, voe, gms_ets, gms_etd);
}

void cc_triples_(double *eh, double *ep, double *v1, double *t1, double *t2, double *v3, double *t3, double *vm,
                 double *voe, double *ve_i, double *ve_j, double *ve_k)
{
    (void)v3; (void)t3; (void)ve_i; (void)ve_j; (void)ve_k;
    legacy_triples(g_d_vvvo, eh, ep, v1, t1, t2, vm, voe, NULL, NULL);
}

/* ------------------------------------------------------------------ per-tuple engine */

/* The per-tuple entries (ijk_tuple_, t1wt3_ijk_, ddcc_t_*_acc_, ddcc_t_ijk_big_, the ijk_gpu_*
 * trio) drive the reference's tuple loop one call at a time.  They share one persistent device
 * workspace, kept across calls while (no, nu) stays the same: operand panels for three occupied
 * slots (the tuple's indices), BX panels and padded VOE blocks for the 3 x 3 ordered slot pairs,
 * and the partial buffers of one tuple's work items.  A call uploads only the slices it is given
 * or that its tuple reads (VE_x or the packed VVVO patch, T2(:,:,:,x), VM(:,:,z,y), VOE(:,:,y,z)),
 * packs those panels, and runs the persistent kernel over its one tuple with the tuple's indices
 * renamed to slots 0..2 (the kernel addresses AX by slot, BX and VOE by slot pair). */
#define TE_SLOTS 3
static struct {
    int64_t no, nu, nup, nb, kp, nblk;
    double *axr, *axc, *bx, *voep, *eh, *ep, *t1, *epart, *v1part, *v1t, *v1, *out;
    double *vvvox, *t2x, *vmb, *voeb, *in, *x3, *v1p;
    int *tup, *blocks;
    unsigned int *counter;
    double *h_ep;
    int ready;
} g_te;

static void te_free(void)
{
    double *d[] = {g_te.axr, g_te.axc, g_te.bx, g_te.voep, g_te.eh, g_te.ep, g_te.t1, g_te.epart, g_te.v1part,
                   g_te.v1t, g_te.v1, g_te.out, g_te.vvvox, g_te.t2x, g_te.vmb, g_te.voeb, g_te.in, g_te.x3,
                   g_te.v1p};
    for (size_t q = 0; q < sizeof(d) / sizeof(d[0]); ++q) ccx_hip_free(d[q]);
    ccx_hip_free(g_te.tup);
    ccx_hip_free(g_te.blocks);
    ccx_hip_free(g_te.counter);
    free(g_te.h_ep);
    memset(&g_te, 0, sizeof(g_te));
}

/* the workspace for (no, nu), allocated on first use and whenever the dimensions change */
static int te_ensure(int64_t no, int64_t nu)
{
    if (no < 1 || nu < 1 || no > 2048 || nu > 4096) return CCX_ERR_ARG;
    if (g_te.ready && g_te.no == no && g_te.nu == nu) return CCX_OK;
    te_free();
    int ndev = 0;
    if (ccx_hip_device_count(&ndev) != 0 || ndev < 1) return CCX_ERR_DEVICE;
    int rc = CCX_OK;
    int *h_blk = NULL;
    const int64_t nup = (nu + TB - 1) / TB * TB, nb = nup / TB, kp = (nu + no + 3) / 4 * 4;
    const int64_t nblk = build_blocks(nb, NULL), S = TE_SLOTS, nu3 = nu * nu * nu;
    g_te.no = no; g_te.nu = nu; g_te.nup = nup; g_te.nb = nb; g_te.kp = kp; g_te.nblk = nblk;
    DALLOC(g_te.axr, S * nb * nb * kp * 256);
    DALLOC(g_te.axc, S * nb * nb * kp * 256);
    DALLOC(g_te.bx, S * S * nb * kp * 16);
    DALLOC(g_te.voep, S * S * nup * nup);
    DALLOC(g_te.eh, S);
    DALLOC(g_te.ep, nup);
    DALLOC(g_te.t1, S * nu);
    DALLOC(g_te.epart, nblk);
    DALLOC(g_te.v1part, nblk * 144);
    DALLOC(g_te.v1t, 3 * nu);
    DALLOC(g_te.v1, S * nu);
    DALLOC(g_te.out, 4);
    DALLOC(g_te.vvvox, S * nu3);
    DALLOC(g_te.t2x, S * nu * nu * no);
    DALLOC(g_te.vmb, S * S * no * nu);
    DALLOC(g_te.voeb, S * S * nu * nu);
    DALLOC(g_te.in, nu3);
    DALLOC(g_te.x3, nu);
    DALLOC(g_te.v1p, 3 * nu);
    DALLOC(g_te.tup, 4);
    DALLOC(g_te.blocks, 4 * nblk);
    DALLOC(g_te.counter, 8);
    g_te.h_ep = (double *)malloc(sizeof(double) * (size_t)nup);
    h_blk = (int *)malloc(sizeof(int) * 4 * (size_t)nblk);
    if (!g_te.h_ep || !h_blk) { rc = CCX_ERR_ALLOC; goto fail; }
    build_blocks(nb, h_blk);
    /* everything a call does not overwrite stays zero and finite */
    if (ccx_hip_zero(g_te.voep, sizeof(double) * (size_t)(S * S * nup * nup), NULL) != 0 ||
        ccx_hip_zero(g_te.eh, sizeof(double) * (size_t)S, NULL) != 0 ||
        ccx_hip_zero(g_te.t1, sizeof(double) * (size_t)(S * nu), NULL) != 0 ||
        ccx_hip_memcpy_h2d(g_te.blocks, h_blk, sizeof(int) * 4 * (size_t)nblk, NULL) != 0) {
        rc = CCX_ERR_DEVICE;
        goto fail;
    }
    for (int64_t a = 0; a < nup; ++a) g_te.h_ep[a] = 1.0e3;
    if (ccx_hip_memcpy_h2d(g_te.ep, g_te.h_ep, sizeof(double) * (size_t)nup, NULL) != 0) { rc = CCX_ERR_DEVICE; goto fail; }
    free(h_blk);
    g_te.ready = 1;
    return CCX_OK;
fail:
    free(h_blk);
    te_free();
    return rc;
}

/* One tuple through the workspace.  x[p]: the occupied index (0-based) at tuple position p; kind
 * 1 (i,j,k), 2 (i,j,j), 3 (i,i,j).  vvvo_src[p]: the cube of position p's index, as VE_x(d,p,q)
 * = VVVO(d,q,p,x) (src_kind 0) or as the distributed packed patch, nutr x nu (src_kind 1);
 * NULL for a position that repeats an earlier index.  t2x[p] = T2(:,:,:,x[p]) (likewise NULL for
 * a repeat).  vm(y,z) returns VM(:,:,x_y,x_z) for tuple positions y,z, voe(y,z) VOE(:,:,x_y,x_z);
 * with voe == NULL no energy is formed.  Energy weights are those of the full set (1, 1/2).
 * etd, v1 (nu,no, +=, columns x) and w (nu^3, W_ijk with W(a,a,a) = 0) are each optional. */
typedef const double *(*te_block_fn)(const void *ud, int y, int z);
static int te_tuple(int64_t no, int64_t nu, const int64_t x[3], int kind, const double *const vvvo_src[3],
                    int src_kind, const double *const t2x[3], te_block_fn vm, te_block_fn voe, const void *ud,
                    const double *eh, const double *ep, double *etd, double *v1, double *w)
{
    int rc = te_ensure(no, nu);
    if (rc != CCX_OK) return rc;
    for (int p = 0; p < 3; ++p)
        if (x[p] < 0 || x[p] >= no) return CCX_ERR_ARG;
    const int64_t nu2 = nu * nu, nu3 = nu2 * nu, nup = g_te.nup, kp = g_te.kp, nb = g_te.nb;
    const int64_t nutr = nu * (nu + 1) / 2;
    /* slots: position p -> slot sl[p]; the first position holding an index owns its slot */
    int sl[3], nslot = 0;
    for (int p = 0; p < 3; ++p) {
        sl[p] = -1;
        for (int q = 0; q < p; ++q)
            if (x[q] == x[p]) sl[p] = sl[q];
        if (sl[p] < 0) sl[p] = nslot++;
    }
    const int64_t panel = nb * nb * kp * 256;
    double h_eh[TE_SLOTS] = {0.0, 0.0, 0.0};
    if (eh)
        for (int p = 0; p < 3; ++p) h_eh[sl[p]] = eh[x[p]];
    for (int p = 0; p < 3; ++p) {
        int first = 1;
        for (int q = 0; q < p; ++q)
            if (x[q] == x[p]) first = 0;
        if (!first) continue;
        if (!vvvo_src[p] || !t2x[p]) return CCX_ERR_ARG;
        double *vx = g_te.vvvox + sl[p] * nu3;
        if (src_kind == 0) {
            if (ccx_hip_memcpy_h2d(g_te.in, vvvo_src[p], sizeof(double) * (size_t)nu3, NULL) != 0 ||
                ccx_hip_ve_to_vvvo(g_te.in, vx, (int)nu, NULL) != 0)
                return CCX_ERR_DEVICE;
        } else {
            if (ccx_hip_memcpy_h2d(g_te.in, vvvo_src[p], sizeof(double) * (size_t)(nutr * nu), NULL) != 0 ||
                ccx_hip_unpack_vvvo(g_te.in, vx, 1, (int)nu, NULL) != 0)
                return CCX_ERR_DEVICE;
        }
        double *tx = g_te.t2x + sl[p] * nu2 * no;
        if (ccx_hip_memcpy_h2d(tx, t2x[p], sizeof(double) * (size_t)(nu2 * no), NULL) != 0 ||
            ccx_hip_pack_ax_one(vx, tx, g_te.axr + sl[p] * panel, g_te.axc + sl[p] * panel, (int)no, (int)nu,
                                (int)nup, (int)kp, NULL) != 0)
            return CCX_ERR_DEVICE;
    }
    /* BX panels and VOE blocks of the ordered position pairs (y,z), y != z, at slot pair (sy,sz) */
    for (int y = 0; y < 3; ++y)
        for (int z = 0; z < 3; ++z) {
            if (y == z) continue;
            const int sy = sl[y], sz = sl[z], sp = sy * TE_SLOTS + sz;
            const double *vmzy = vm(ud, z, y), *voeyz = voe ? voe(ud, y, z) : NULL;
            if (!vmzy || (voe && !voeyz)) return CCX_ERR_ARG;
            double *vmd = g_te.vmb + sp * no * nu, *voed = g_te.voeb + sp * nu2;
            if (ccx_hip_memcpy_h2d(vmd, vmzy, sizeof(double) * (size_t)(no * nu), NULL) != 0 ||
                (voeyz && ccx_hip_memcpy_h2d(voed, voeyz, sizeof(double) * (size_t)nu2, NULL) != 0))
                return CCX_ERR_DEVICE;
            /* T2(:,:,x_y,x_z) sits in the T2(:,:,:,x_z) slab of slot sz */
            const double *t2yz = g_te.t2x + sz * nu2 * no + x[y] * nu2;
            if (ccx_hip_pack_pair(t2yz, vmd, voeyz ? voed : NULL, g_te.bx + sp * nb * kp * 16,
                                  g_te.voep + sp * nup * nup, (int)no, (int)nu, (int)nup, (int)kp, NULL) != 0)
                return CCX_ERR_DEVICE;
        }
    if (ep) {
        memcpy(g_te.h_ep, ep, sizeof(double) * (size_t)nu);
        if (ccx_hip_memcpy_h2d(g_te.ep, g_te.h_ep, sizeof(double) * (size_t)nup, NULL) != 0) return CCX_ERR_DEVICE;
    }
    const int tup[4] = {sl[0], sl[1], sl[2], kind};
    if (ccx_hip_memcpy_h2d(g_te.eh, h_eh, sizeof(h_eh), NULL) != 0 ||
        ccx_hip_memcpy_h2d(g_te.tup, tup, sizeof(tup), NULL) != 0)
        return CCX_ERR_DEVICE;
    int cus = 256;
    if (ccx_hip_cu_count(&cus) != 0 || cus < 1) cus = 256;
    const int grid = (int)(g_te.nblk < cus ? g_te.nblk : cus);
    if (ccx_hip_triples(0, g_te.axr, g_te.axc, g_te.bx, g_te.voep, g_te.eh, g_te.ep, g_te.tup, g_te.blocks,
                        g_te.epart, g_te.v1part, w ? g_te.in : NULL, g_te.counter, TE_SLOTS, (int)nu, (int)nup,
                        (int)kp, 1, (int)g_te.nblk, grid, NULL, 0, 1.0, NULL) != 0)
        return CCX_ERR_DEVICE;
    if (w && ccx_hip_memcpy_d2h(w, g_te.in, sizeof(double) * (size_t)nu3, NULL) != 0) return CCX_ERR_DEVICE;
    if (w)
        for (int64_t a = 0; a < nu; ++a) w[a + nu * (a + nu * a)] = 0.0; /* zerot3 */
    if (voe && (etd || v1)) {
        double out[2];
        double *hv = (double *)malloc(sizeof(double) * (size_t)(TE_SLOTS * nu));
        if (!hv) return CCX_ERR_ALLOC;
        if (ccx_hip_finalize(g_te.epart, g_te.v1part, g_te.tup, g_te.blocks, g_te.t1, g_te.v1t, g_te.v1, g_te.out,
                             TE_SLOTS, (int)nu, 1, (int)g_te.nblk, NULL) != 0 ||
            ccx_hip_memcpy_d2h(out, g_te.out, sizeof(out), NULL) != 0 ||
            ccx_hip_memcpy_d2h(hv, g_te.v1, sizeof(double) * (size_t)(TE_SLOTS * nu), NULL) != 0) {
            free(hv);
            return CCX_ERR_DEVICE;
        }
        if (etd) *etd += out[0];
        if (v1)   /* slot column s -> occupied column of the first position holding it */
            for (int s2 = 0; s2 < nslot; ++s2) {
                int64_t col = -1;
                for (int p = 0; p < 3 && col < 0; ++p)
                    if (sl[p] == s2) col = x[p];
                for (int64_t a = 0; a < nu; ++a) v1[a + nu * col] += hv[a + nu * s2];
            }
        free(hv);
    }
    return CCX_OK;
}

/* block lookups for the two argument styles: full VM / VOE arrays, or the six pair slices */
struct te_full { const double *vm, *voe; int64_t no, nu; int64_t x[3]; };
static const double *te_full_vm(const void *ud, int y, int z)
{
    const struct te_full *f = (const struct te_full *)ud;   /* VM(:,:,x_y,x_z) */
    return f->vm + f->no * f->nu * (f->x[y] + f->no * f->x[z]);
}
static const double *te_full_voe(const void *ud, int y, int z)
{
    const struct te_full *f = (const struct te_full *)ud;   /* VOE(:,:,x_y,x_z) */
    return f->voe + f->nu * f->nu * (f->x[y] + f->no * f->x[z]);
}
struct te_pairs { const double *blk[3][3]; };               /* blk[y][z] = (:,:,x_y,x_z) */
static const double *te_pair_blk(const void *ud, int y, int z) { return ((const struct te_pairs *)ud)->blk[y][z]; }

/* one tuple from the full arrays the driver passes: v1 (nu,no) +=, etd +=, W into v3 when non-null */
static int te_full_tuple(int64_t no, int64_t nu, const int64_t x[3], int kind, const double *const src[3], int src_kind,
                         const double *t2, const double *vm, const double *voe, const double *eh, const double *ep,
                         double *etd, double *v1, double *v3)
{
    struct te_full f = {vm, voe, no, nu, {x[0], x[1], x[2]}};
    const double *t2x[3];
    for (int p = 0; p < 3; ++p) t2x[p] = t2 + nu * nu * no * x[p];
    return te_tuple(no, nu, x, kind, src, src_kind, t2x, te_full_vm, te_full_voe, &f, eh, ep, etd, v1, v3);
}

void ijk_tuple_(int64_t *i, int64_t *j, int64_t *k, const double *t2_i, const double *t2_j, const double *t2_k,
                const double *vm_ij, const double *vm_ji, const double *vm_ik, const double *vm_ki,
                const double *vm_jk, const double *vm_kj, const double *ve_i, const double *ve_j,
                const double *ve_k, double *v3)
{
    if (!i || !j || !k || !t2_i || !t2_j || !t2_k || !vm_ij || !vm_ji || !vm_ik || !vm_ki || !vm_jk || !vm_kj ||
        !ve_i || !ve_j || !ve_k || !v3) { g_status = CCX_ERR_ARG; return; }
    if (g_no < 1 || g_nu < 1) { g_status = CCX_ERR_STATE; return; }
    const int64_t x[3] = {*i - 1, *j - 1, *k - 1};
    if (x[0] == x[1] || x[1] == x[2] || x[0] == x[2]) { g_status = CCX_ERR_ARG; return; }
    struct te_pairs pr = {{{NULL, vm_ij, vm_ik}, {vm_ji, NULL, vm_jk}, {vm_ki, vm_kj, NULL}}};
    const double *src[3] = {ve_i, ve_j, ve_k}, *t2x[3] = {t2_i, t2_j, t2_k};
    g_status = te_tuple(g_no, g_nu, x, 1, src, 0, t2x, te_pair_blk, NULL, &pr, NULL, NULL, NULL, NULL, v3);
}

/* E[T] and v1 of one tuple (0-based i,j,k) from its W cube on the GPU (t1wt3_kernel):
 * etd += efac * x3, v1(:,i/j/k) += vfac * (row sums); v1 is (nu,no). */
static int t1wt3_host(int64_t no, int64_t nu, int64_t ii, int64_t jj, int64_t kk, const double *v3,
                      const double *voe6[6], double *v1, const double *eh, const double *ep, double efac,
                      double vfac)
{
    if (ii < 0 || jj < 0 || kk < 0 || ii >= no || jj >= no || kk >= no) return CCX_ERR_ARG;
    int rc = te_ensure(no, nu);
    if (rc != CCX_OK) return rc;
    const size_t nu2 = (size_t)(nu * nu), nu3 = nu2 * (size_t)nu;
    double *x3 = (double *)malloc(sizeof(double) * (size_t)nu), *v1p = (double *)malloc(sizeof(double) * 3 * (size_t)nu);
    if (!x3 || !v1p) { rc = CCX_ERR_ALLOC; goto done; }
    /* W in the staging cube, the six VOE slices in the first six pair blocks, eps_virtual in ep */
    if (ccx_hip_memcpy_h2d(g_te.in, v3, sizeof(double) * nu3, NULL) != 0 ||
        ccx_hip_memcpy_h2d(g_te.ep, ep, sizeof(double) * (size_t)nu, NULL) != 0) { rc = CCX_ERR_DEVICE; goto done; }
    for (int q = 0; q < 6; ++q)
        if (ccx_hip_memcpy_h2d(g_te.voeb + q * nu2, voe6[q], sizeof(double) * nu2, NULL) != 0) { rc = CCX_ERR_DEVICE; goto done; }
    if (ccx_hip_t1wt3(g_te.in, g_te.voeb, eh[ii] + eh[jj] + eh[kk], g_te.ep, (int)nu, g_te.x3, g_te.v1p, NULL) != 0 ||
        ccx_hip_memcpy_d2h(x3, g_te.x3, sizeof(double) * (size_t)nu, NULL) != 0 ||
        ccx_hip_memcpy_d2h(v1p, g_te.v1p, sizeof(double) * 3 * (size_t)nu, NULL) != 0) { rc = CCX_ERR_DEVICE; goto done; }
    {
        double e = 0.0;
        for (int64_t a = 0; a < nu; ++a) e += x3[a];
        g_etd += efac * e;
        const int64_t rows[3] = {ii, jj, kk};
        for (int r = 0; r < 3; ++r)
            for (int64_t a = 0; a < nu; ++a) v1[a + nu * rows[r]] += vfac * v1p[3 * a + r];
    }
done:   /* only ep[0..nu) was overwritten: the padding the tuple kernel reads stays finite */
    free(x3); free(v1p);
    return rc;
}

void t1wt3_ijk_(int64_t *i, int64_t *j, int64_t *k, int64_t *no, int64_t *nu, const double *v3,
                const double *voe_ij, const double *voe_ji, const double *voe_ik, const double *voe_ki,
                const double *voe_jk, const double *voe_kj, double *t1, const double *eh, const double *ep)
{
    if (!i || !j || !k || !no || !nu || !v3 || !voe_ij || !voe_ji || !voe_ik || !voe_ki || !voe_jk || !voe_kj ||
        !t1 || !eh || !ep) { g_status = CCX_ERR_ARG; return; }
    const double *voe6[6] = {voe_ij, voe_ji, voe_ik, voe_ki, voe_jk, voe_kj};
    /* the reference halves the energy of a tuple with equal indices (t1wt3.f90:98-102) */
    const double cf = (*i == *j || *j == *k) ? 0.5 : 1.0;
    g_status = t1wt3_host(*no, *nu, *i - 1, *j - 1, *k - 1, v3, voe6, t1, eh, ep, cf, 1.0);
}

/* ddcc_t_ijk_acc(no,nu,i,j,k,t1,t2,vm,v3,voe,eh,ep,ve_i,ve_j,ve_k): src/triples-ijk/ddcc_t_ijk_big.f:1-21.
 * The argument named t1 is the v1 accumulator (nu,no), as the driver passes it. */
void ddcc_t_ijk_acc_(int64_t *no, int64_t *nu, int64_t *i, int64_t *j, int64_t *k, double *t1, double *t2,
                     double *vm, double *v3, double *voe, double *eh, double *ep, double *ve_i, double *ve_j,
                     double *ve_k)
{
    if (!no || !nu || !i || !j || !k || !t1 || !t2 || !vm || !v3 || !voe || !eh || !ep || !ve_i || !ve_j || !ve_k) {
        g_status = CCX_ERR_ARG;
        return;
    }
    const int64_t x[3] = {*i - 1, *j - 1, *k - 1};
    if (x[0] == x[1] || x[1] == x[2] || x[0] == x[2]) { g_status = CCX_ERR_ARG; return; }
    const double *src[3] = {ve_i, ve_j, ve_k};
    g_status = te_full_tuple(*no, *nu, x, 1, src, 0, t2, vm, voe, eh, ep, &g_etd, t1, v3);
}

/* ddcc_t_ijj_acc(no,nu,i,j,t1,t2,vm,v3,t3,voe,o1,eh,ep,temp,ve_i,ve_j) and the (i,i,j)
 * twin: src/triples-ijk/ddcc_t_ijj_iij_big.f:18-94.  t1 is the v1 accumulator; t3, o1 and
 * temp are the reference's scratch and T1 copy, not read here.  v3 receives W (the
 * reference leaves it in a transformed scratch state). */
void ddcc_t_ijj_acc_(int64_t *no, int64_t *nu, int64_t *i, int64_t *j, double *t1, double *t2, double *vm,
                     double *v3, double *t3, double *voe, double *o1, double *eh, double *ep, double *temp,
                     double *ve_i, double *ve_j)
{
    if (!no || !nu || !i || !j || !t1 || !t2 || !vm || !v3 || !voe || !eh || !ep || !ve_i || !ve_j) {
        g_status = CCX_ERR_ARG;
        return;
    }
    (void)t3; (void)o1; (void)temp;
    const int64_t x[3] = {*i - 1, *j - 1, *j - 1};
    if (x[0] == x[1]) { g_status = CCX_ERR_ARG; return; }
    const double *src[3] = {ve_i, ve_j, NULL};
    g_status = te_full_tuple(*no, *nu, x, 2, src, 0, t2, vm, voe, eh, ep, &g_etd, t1, v3);
}

void ddcc_t_iij_acc_(int64_t *no, int64_t *nu, int64_t *i, int64_t *j, double *t1, double *t2, double *vm,
                     double *v3, double *t3, double *voe, double *o1, double *eh, double *ep, double *temp,
                     double *ve_i, double *ve_j)
{
    if (!no || !nu || !i || !j || !t1 || !t2 || !vm || !v3 || !voe || !eh || !ep || !ve_i || !ve_j) {
        g_status = CCX_ERR_ARG;
        return;
    }
    (void)t3; (void)o1; (void)temp;
    const int64_t x[3] = {*i - 1, *i - 1, *j - 1};
    if (x[0] == x[2]) { g_status = CCX_ERR_ARG; return; }
    const double *src[3] = {ve_i, NULL, ve_j};
    g_status = te_full_tuple(*no, *nu, x, 3, src, 0, t2, vm, voe, eh, ep, &g_etd, t1, v3);
}

/* ------------------------------------------------------------------ permutations (host) */

/* new(x,y,z) = old(s(x,y,z)); id 1 (x,z,y), 2 (y,x,z), 3 (z,y,x), 4 (y,z,x), 5 (z,x,y) */
static int cube_perm(double *v, int64_t n, int id)
{
    if (!v || n < 0 || id < 1 || id > 5) return CCX_ERR_ARG;
    const size_t n3 = (size_t)(n * n * n);
    double *old = (double *)malloc(sizeof(double) * (n3 ? n3 : 1));
    if (!old) return CCX_ERR_ALLOC;
    memcpy(old, v, sizeof(double) * n3);
    for (int64_t z = 0; z < n; ++z)
        for (int64_t y = 0; y < n; ++y)
            for (int64_t x = 0; x < n; ++x) {
                int64_t sx, sy, sz;
                switch (id) {
                case 1: sx = x; sy = z; sz = y; break;
                case 2: sx = y; sy = x; sz = z; break;
                case 3: sx = z; sy = y; sz = x; break;
                case 4: sx = y; sy = z; sz = x; break;
                default: sx = z; sy = x; sz = y; break;
                }
                v[x + n * (y + n * z)] = old[sx + n * (sy + n * sz)];
            }
    free(old);
    return CCX_OK;
}

void trant3_1_(int64_t *n, double *v) { g_status = n ? cube_perm(v, *n, 1) : CCX_ERR_ARG; }
void trant3_3_(int64_t *n, double *v) { g_status = n ? cube_perm(v, *n, 3) : CCX_ERR_ARG; }
void trant3_4_(int64_t *n, double *v) { g_status = n ? cube_perm(v, *n, 4) : CCX_ERR_ARG; }
void trant3_5_(int64_t *n, double *v) { g_status = n ? cube_perm(v, *n, 5) : CCX_ERR_ARG; }
void trant3_acc_(double *v, int64_t *nu, int64_t *id)
{
    g_status = (nu && id) ? cube_perm(v, *nu, (int)*id) : CCX_ERR_ARG;
}
void trant3_smp_(double *v, int64_t *nu, int64_t *id)
{
    g_status = (nu && id) ? cube_perm(v, *nu, (int)*id) : CCX_ERR_ARG;
}

void tranmd_23_(double *a, int64_t *n1, int64_t *n2, int64_t *n3, int64_t *n4)
{
    if (!a || !n1 || !n2 || !n3 || !n4 || *n2 != *n3 || *n1 < 0 || *n2 < 0 || *n4 < 0) {
        g_status = CCX_ERR_ARG;
        return;
    }
    const int64_t N1 = *n1, N2 = *n2, N4 = *n4;
    for (int64_t l = 0; l < N4; ++l)
        for (int64_t j = 0; j < N2; ++j)
            for (int64_t k = 0; k < j; ++k)
                for (int64_t i = 0; i < N1; ++i) {
                    double *p = a + i + N1 * (j + N2 * (k + N2 * l));
                    double *q = a + i + N1 * (k + N2 * (j + N2 * l));
                    double x = *p;
                    *p = *q;
                    *q = x;
                }
    g_status = CCX_OK;
}

/* trant3_1_async(acc_sync,n,v), trant3_4_async(acc_sync,n,v): src/triples-ijk/ijk_tuple_acc.f90:109-153.
 * The queue id selects an OpenACC queue in the reference; these host permutations finish
 * before returning, so it is not read. */
void trant3_1_async_(int64_t *q, int64_t *n, double *v)
{
    (void)q;
    g_status = n ? cube_perm(v, *n, 1) : CCX_ERR_ARG;
}
void trant3_4_async_(int64_t *q, int64_t *n, double *v)
{
    (void)q;
    g_status = n ? cube_perm(v, *n, 4) : CCX_ERR_ARG;
}

/* tranmd_23_acc(acc_sync,a,n1,n2,n3,n4): src/common/common_cc.f90:203-237, tranmd_23 on a queue */
void tranmd_23_acc_(int64_t *q, double *a, int64_t *n1, int64_t *n2, int64_t *n3, int64_t *n4)
{
    (void)q;
    tranmd_23_(a, n1, n2, n3, n4);
}

/* TRANMD_SMP(a,n1,n2,n3,n4,ij): src/triples-ijk/common_functions.f:141-421.  In-place index
 * permutation of a(n1,n2,n3,n4), new(x1,x2,x3,x4) = old(source), by code:
 *   12 (x2,x1,x3,x4)  13 (x3,x2,x1,x4)  14 (x4,x2,x3,x1)  23 (x1,x3,x2,x4)  24 (x1,x4,x3,x2)
 *   34 (x1,x2,x4,x3)  231 (x2,x3,x1,x4)  312 (x3,x1,x2,x4)  341 (x3,x2,x4,x1)  413 (x4,x2,x1,x3)
 *   1234 (x4,x3,x2,x1)
 * Each needs the permuted extents equal (the reference's loops assume it); other extents
 * are rejected with CCX_ERR_ARG and the array left as it was.  Any other code is a no-op,
 * as in the reference. */
void tranmd_smp_(double *a, int64_t *n1, int64_t *n2, int64_t *n3, int64_t *n4, int64_t *ij)
{
    if (!a || !n1 || !n2 || !n3 || !n4 || !ij || *n1 < 0 || *n2 < 0 || *n3 < 0 || *n4 < 0) {
        g_status = CCX_ERR_ARG;
        return;
    }
    static const struct { int code; int src[4]; } tab[] = {
        {12, {1, 0, 2, 3}},  {13, {2, 1, 0, 3}},  {14, {3, 1, 2, 0}},  {23, {0, 2, 1, 3}},
        {24, {0, 3, 2, 1}},  {34, {0, 1, 3, 2}},  {231, {1, 2, 0, 3}}, {312, {2, 0, 1, 3}},
        {341, {2, 1, 3, 0}}, {413, {3, 1, 0, 2}}, {1234, {3, 2, 1, 0}},
    };
    int e = -1;
    for (int q = 0; q < (int)(sizeof(tab) / sizeof(tab[0])); ++q)
        if (tab[q].code == *ij) e = q;
    if (e < 0) { g_status = CCX_OK; return; }
    const int64_t n[4] = {*n1, *n2, *n3, *n4};
    const int *s = tab[e].src;
    for (int d = 0; d < 4; ++d)
        if (n[s[d]] != n[d]) { g_status = CCX_ERR_ARG; return; }
    const size_t total = (size_t)(n[0] * n[1] * n[2] * n[3]);
    double *old = (double *)malloc(sizeof(double) * (total ? total : 1));
    if (!old) { g_status = CCX_ERR_ALLOC; return; }
    memcpy(old, a, sizeof(double) * total);
    const int64_t str[4] = {1, n[0], n[0] * n[1], n[0] * n[1] * n[2]};
    int64_t x[4];
    for (x[3] = 0; x[3] < n[3]; ++x[3])
        for (x[2] = 0; x[2] < n[2]; ++x[2])
            for (x[1] = 0; x[1] < n[1]; ++x[1])
                for (x[0] = 0; x[0] < n[0]; ++x[0]) {
                    const int64_t dst = x[0] + str[1] * x[1] + str[2] * x[2] + str[3] * x[3];
                    const int64_t src = x[s[0]] + str[1] * x[s[1]] + str[2] * x[s[2]] + str[3] * x[s[3]];
                    a[dst] = old[src];
                }
    free(old);
    g_status = CCX_OK;
}

/* DDCC_T_IJK_BIG(no,nu,i,j,k,t1,t2,vm,v3,t3,voe,o1,tmp,eh,ep,ve_i,ve_j,ve_k), the triples-abc
 * per-tuple routine (src/triples-abc/ddcc_t_ijk_big.f:1-32, called with t1 = v1 and o1 = T1 at
 * src/triples-abc/triples.f90:255).  Its T3SQUA_SMP energy and DRT1WT3IJK_SMP v1 update are
 * the same sums as t1wt3_ijk's (weight 1 for i > j > k), so it shares ddcc_t_ijk_acc_'s
 * implementation; v3 receives W (the reference leaves W/D, permuted), t3, o1 and tmp are
 * not read. */
void ddcc_t_ijk_big_(int64_t *no, int64_t *nu, int64_t *i, int64_t *j, int64_t *k, double *t1, double *t2,
                     double *vm, double *v3, double *t3, double *voe, double *o1, double *tmp, double *eh, double *ep,
                     double *ve_i, double *ve_j, double *ve_k)
{
    (void)t3; (void)o1; (void)tmp;
    ddcc_t_ijk_acc_(no, nu, i, j, k, t1, t2, vm, v3, voe, eh, ep, ve_i, ve_j, ve_k);
}

/* ------------------------------------------------------------------ the CUDA driver trio */

/* ijk_gpu_init(no,nu,eh,ep,v1,t2,vm,voe) / ijk_gpu_driver(nu,no,i,j,k,ve_i,ve_j,ve_k,t1,eh,ep,etd)
 * / ijk_gpu_finalize(no,nu,etd): the reference's CUDA tuple path (call sites
 * src/triples-ijk/triples_cpu.F90:363,465,497 and ddcc_t_ijk_big_cuda.f:1-16; its sources are
 * not in the tree).  init keeps the caller's arrays and clears the accumulators; driver takes
 * each occupied index's raw VVVO patch as the distributed array holds it (nutr x nu, column c
 * = the lower-packed (p,q) block of VVVO(:,:,c,x)) and does the expansion and index swap
 * itself, then adds the tuple's E[T] and v1 contributions to the accumulators; finalize adds
 * them to *etd and to the v1 passed to init.  The driver's own t1/etd arguments are not read. */
static struct {
    int active;
    int64_t no, nu;
    double *v1;
    const double *eh, *ep, *t2, *vm, *voe;
    double *acc_v1;
    double etd;
} g_gpu;

static void gpu_trio_free(void)
{
    free(g_gpu.acc_v1);
    memset(&g_gpu, 0, sizeof(g_gpu));
}

void ijk_gpu_init_(int64_t *no, int64_t *nu, double *eh, double *ep, double *v1, double *t2, double *vm, double *voe)
{
    if (!no || !nu || *no < 1 || *nu < 1 || !eh || !ep || !v1 || !t2 || !vm || !voe) { g_status = CCX_ERR_ARG; return; }
    gpu_trio_free();
    g_gpu.acc_v1 = (double *)calloc((size_t)(*no * *nu), sizeof(double));
    if (!g_gpu.acc_v1) { g_status = CCX_ERR_ALLOC; return; }
    int rc = te_ensure(*no, *nu);
    if (rc != CCX_OK) { gpu_trio_free(); g_status = rc; return; }
    g_gpu.no = *no; g_gpu.nu = *nu;
    g_gpu.v1 = v1; g_gpu.eh = eh; g_gpu.ep = ep; g_gpu.t2 = t2; g_gpu.vm = vm; g_gpu.voe = voe;
    g_gpu.etd = 0.0;
    g_gpu.active = 1;
    g_status = CCX_OK;
}

void ijk_gpu_driver_(int64_t *nu, int64_t *no, int64_t *i, int64_t *j, int64_t *k, double *ve_i, double *ve_j,
                     double *ve_k, double *t1, double *eh, double *ep, double *etd)
{
    (void)t1; (void)eh; (void)ep; (void)etd;
    if (!g_gpu.active) { g_status = CCX_ERR_STATE; return; }
    if (!nu || !no || !i || !j || !k || !ve_i || !ve_j || !ve_k || *nu != g_gpu.nu || *no != g_gpu.no) {
        g_status = CCX_ERR_ARG;
        return;
    }
    const int64_t x[3] = {*i - 1, *j - 1, *k - 1};
    if (x[0] == x[1] || x[1] == x[2] || x[0] == x[2]) { g_status = CCX_ERR_ARG; return; }
    /* the raw packed patches: the device expands them and swaps the index pair */
    const double *src[3] = {ve_i, ve_j, ve_k};
    g_status = te_full_tuple(g_gpu.no, g_gpu.nu, x, 1, src, 1, g_gpu.t2, g_gpu.vm, g_gpu.voe, g_gpu.eh, g_gpu.ep,
                             &g_gpu.etd, g_gpu.acc_v1, NULL);
}

void ijk_gpu_finalize_(int64_t *no, int64_t *nu, double *etd)
{
    if (!g_gpu.active) { g_status = CCX_ERR_STATE; return; }
    if (!no || !nu || !etd || *no != g_gpu.no || *nu != g_gpu.nu) { g_status = CCX_ERR_ARG; return; }
    *etd += g_gpu.etd;
    for (int64_t q = 0; q < g_gpu.no * g_gpu.nu; ++q) g_gpu.v1[q] += g_gpu.acc_v1[q];
    gpu_trio_free();
    g_status = CCX_OK;
}

/* ------------------------------------------------------------------ triples.restart */

/* The reference's input file (src/triples-ijk/program.F90:44-46,195-218): Fortran unformatted
 * sequential records, each framed by a 4-byte length before and after, int64 integers:
 * no; nu; eh(no); ep(nu); t1(no*nu); t2(no^2 nu^2); vm(no^3 nu); voe(no^2 nu^2); then no*nu
 * records of nutr doubles, record j = column j = c + nu*x of the lower-packed VVVO.
 * Records above 2^31-1 bytes (not representable with 4-byte markers) are rejected. */
static int rec_read(FILE *f, void *buf, size_t bytes)
{
    int32_t m0 = 0, m1 = 0;
    if (bytes > (size_t)INT32_MAX) return CCX_ERR_IO;
    if (fread(&m0, 4, 1, f) != 1 || (size_t)m0 != bytes) return CCX_ERR_IO;
    if (bytes && fread(buf, 1, bytes, f) != bytes) return CCX_ERR_IO;
    if (fread(&m1, 4, 1, f) != 1 || m1 != m0) return CCX_ERR_IO;
    return CCX_OK;
}

static int rec_write(FILE *f, const void *buf, size_t bytes)
{
    if (bytes > (size_t)INT32_MAX) return CCX_ERR_IO;
    const int32_t m = (int32_t)bytes;
    if (fwrite(&m, 4, 1, f) != 1 || (bytes && fwrite(buf, 1, bytes, f) != bytes) || fwrite(&m, 4, 1, f) != 1)
        return CCX_ERR_IO;
    return CCX_OK;
}

int ccx_restart_dims(const char *path, int64_t *no, int64_t *nu)
{
    if (!path || !no || !nu) return CCX_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return CCX_ERR_IO;
    int rc = rec_read(f, no, 8);
    if (rc == CCX_OK) rc = rec_read(f, nu, 8);
    fclose(f);
    if (rc == CCX_OK && (*no < 1 || *nu < 1)) rc = CCX_ERR_IO;
    return rc;
}

int ccx_read_restart(const char *path, int64_t no, int64_t nu, double *eh, double *ep, double *t1, double *t2,
                     double *vm, double *voe, double *vvvo_packed)
{
    if (!path || no < 1 || nu < 1 || !eh || !ep || !t1 || !t2 || !vm || !voe || !vvvo_packed) return CCX_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return CCX_ERR_IO;
    int64_t fno = 0, fnu = 0;
    const size_t nou = (size_t)(no * nu), no2u2 = (size_t)(no * no * nu * nu), nutr = (size_t)(nu * (nu + 1) / 2);
    int rc = rec_read(f, &fno, 8);
    if (rc == CCX_OK) rc = rec_read(f, &fnu, 8);
    if (rc == CCX_OK && (fno != no || fnu != nu)) rc = CCX_ERR_ARG;
    if (rc == CCX_OK) rc = rec_read(f, eh, 8 * (size_t)no);
    if (rc == CCX_OK) rc = rec_read(f, ep, 8 * (size_t)nu);
    if (rc == CCX_OK) rc = rec_read(f, t1, 8 * nou);
    if (rc == CCX_OK) rc = rec_read(f, t2, 8 * no2u2);
    if (rc == CCX_OK) rc = rec_read(f, vm, 8 * (size_t)(no * no * no * nu));
    if (rc == CCX_OK) rc = rec_read(f, voe, 8 * no2u2);
    for (size_t jcol = 0; rc == CCX_OK && jcol < nou; ++jcol) rc = rec_read(f, vvvo_packed + nutr * jcol, 8 * nutr);
    fclose(f);
    return rc;
}

int ccx_write_restart(const char *path, int64_t no, int64_t nu, const double *eh, const double *ep, const double *t1,
                      const double *t2, const double *vm, const double *voe, const double *vvvo, int vvvo_layout)
{
    if (!path || no < 1 || nu < 1 || !eh || !ep || !t1 || !t2 || !vm || !voe || !vvvo ||
        (vvvo_layout != 0 && vvvo_layout != 1))
        return CCX_ERR_ARG;
    const size_t nou = (size_t)(no * nu), no2u2 = (size_t)(no * no * nu * nu), nutr = (size_t)(nu * (nu + 1) / 2);
    double *col = (double *)malloc(sizeof(double) * nutr);
    if (!col) return CCX_ERR_ALLOC;
    FILE *f = fopen(path, "wb");
    if (!f) { free(col); return CCX_ERR_IO; }
    int rc = rec_write(f, &no, 8);
    if (rc == CCX_OK) rc = rec_write(f, &nu, 8);
    if (rc == CCX_OK) rc = rec_write(f, eh, 8 * (size_t)no);
    if (rc == CCX_OK) rc = rec_write(f, ep, 8 * (size_t)nu);
    if (rc == CCX_OK) rc = rec_write(f, t1, 8 * nou);
    if (rc == CCX_OK) rc = rec_write(f, t2, 8 * no2u2);
    if (rc == CCX_OK) rc = rec_write(f, vm, 8 * (size_t)(no * no * no * nu));
    if (rc == CCX_OK) rc = rec_write(f, voe, 8 * no2u2);
    for (size_t jcol = 0; rc == CCX_OK && jcol < nou; ++jcol) {
        const double *src = vvvo_layout == 1 ? vvvo + nutr * jcol : NULL;
        if (vvvo_layout == 0) {   /* lower-packed (p >= q) rows of VVVO(:,:,c,x), p outer */
            const double *blk = vvvo + (size_t)(nu * nu) * jcol;
            size_t r = 0;
            for (int64_t p = 0; p < nu; ++p)
                for (int64_t q = 0; q <= p; ++q) col[r++] = blk[p + nu * q];
            src = col;
        }
        rc = rec_write(f, src, 8 * nutr);
    }
    if (fclose(f) != 0 && rc == CCX_OK) rc = CCX_ERR_IO;
    free(col);
    return rc;
}

int ccx_load_restart(ccx_ctx *c, const char *path)
{
    if (!c || !path) return CCX_ERR_ARG;
    int64_t no = 0, nu = 0;
    int rc = ccx_restart_dims(path, &no, &nu);
    if (rc != CCX_OK) return rc;
    if (no != c->no || nu != c->nu) return CCX_ERR_ARG;
    const size_t nou = (size_t)(no * nu), no2u2 = (size_t)(no * no * nu * nu), nutr = (size_t)(nu * (nu + 1) / 2);
    double *eh = (double *)malloc(8 * (size_t)no), *ep = (double *)malloc(8 * (size_t)nu);
    double *t1 = (double *)malloc(8 * nou), *t2 = (double *)malloc(8 * no2u2);
    double *vm = (double *)malloc(8 * (size_t)(no * no * no * nu)), *voe = (double *)malloc(8 * no2u2);
    double *vv = (double *)malloc(8 * nutr * nou);
    if (!eh || !ep || !t1 || !t2 || !vm || !voe || !vv) rc = CCX_ERR_ALLOC;
    if (rc == CCX_OK) rc = ccx_read_restart(path, no, nu, eh, ep, t1, t2, vm, voe, vv);
    if (rc == CCX_OK) rc = ccx_load_host(c, eh, ep, t1, t2, vm, voe, vv, 1);
    free(eh); free(ep); free(t1); free(t2); free(vm); free(voe); free(vv);
    return rc;
}
