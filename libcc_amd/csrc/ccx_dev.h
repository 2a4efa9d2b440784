/* Internal: the thin C ABI between the C host layer (ccx_host.c) and the HIP kernels
 * (ccx_kernels.hip).  Every function returns 0 on success.  Streams and events are
 * opaque void* (hipStream_t / hipEvent_t). */
#ifndef CCX_DEV_H
#define CCX_DEV_H

#include <stddef.h>

int ccx_hip_device_count(int *n);
int ccx_hip_malloc(void **p, size_t bytes);
int ccx_hip_free(void *p);
int ccx_hip_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int ccx_hip_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int ccx_hip_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int ccx_hip_sync(void *stream);
int ccx_hip_cu_count(int *n);
const char *ccx_hip_last_error(void);
int ccx_hip_lds_bytes(void);

int ccx_hip_unpack_vvvo(const double *packed, double *full, int no, int nu, void *stream);
int ccx_hip_nonfinite(const double *const *p, const size_t *n, int narr, unsigned int *d_flag, int *nonfinite,
                      void *stream);
int ccx_hip_pack(int form, const double *vvvo, const double *t2, const double *vm, const double *voe, void *axr,
                 void *axc, void *bx, double *voep, int no, int nu, int nup, int kp, void *stream);
int ccx_hip_pack_ax_one(const double *vvvo_x, const double *t2_x, double *axr_x, double *axc_x, int no, int nu,
                        int nup, int kp, void *stream);
int ccx_hip_pack_pair(const double *t2_yz, const double *vm_zy, const double *voe_yz, double *bx_yz,
                      double *voep_yz, int no, int nu, int nup, int kp, void *stream);
int ccx_hip_ve_to_vvvo(const double *ve, double *vvvo_x, int nu, void *stream);
int ccx_hip_triples(int form, const void *axr, const void *axc, const void *bx, const double *voep,
                    const double *eh, const double *ep, const int *tuples, const int *blocks, double *e_part,
                    double *v1_part, double *wout, unsigned int *counter, int no, int nu, int nup, int kp,
                    int ntup, int nblk, int grid, const void *const *split, int kps, double wscale, void *stream);
int ccx_hip_form_kpad(int form);
int ccx_hip_split_elem_bytes(int form);
int ccx_hip_pack_split(int form, const double *axr, const double *axc, const double *bx, void *const *sl,
                       unsigned long long *dmax, int no, int nb, int kp, int kps, double *wscale, void *stream);
int ccx_hip_pack_f32v(const double *axr, const double *axc, const double *bx, float *axr32, float *axc32,
                      float *bx32, int no, int nb, int kp, int kpf, void *stream);
int ccx_hip_finalize(const double *e_part, const double *v1_part, const int *tuples, const int *blocks,
                     const double *t1, double *v1t, double *v1, double *out, int no, int nu, int ntup, int nblk,
                     void *stream);

int ccx_hip_t1wt3(const double *w, const double *voe6, double dijk, const double *ep, int nu, double *x3,
                  double *v1p, void *stream);

/* virtual-triple view (ccx_abc.hip): trip holds (a, b, c, kind) per triple, 0-based */
int ccx_hip_w_abc(const void *axr, const void *bx, const int *trip, int ntrip, int no, int nu, int nb, int kp,
                  double *w, void *stream);
int ccx_hip_x_abc(const double *w, const int *trip, int ntrip, int no, const double *eh, const double *ep,
                  double *x, void *stream);

/* adjoint (ccx_adjoint.hip) */
int ccx_hip_adj_lds_bytes(void);
int ccx_hip_adj_items(const void *axr, const void *axc, const void *bx, const double *voep, const double *eh,
                      const double *ep, const double *t1, const int *tuples, const int *blocks, double *e_part,
                      double *v1_part, double *wbar, double *slots, double *wstage, unsigned int *counter, int no,
                      int nu, int nup, int kp, int ntup, int nblk, int t0, int tcount, int grid,
                      int form, const void *const *w32, int kps, double wscale, void *stream);
size_t ccx_hip_adj_stage_doubles(void);
int ccx_hip_gax(const double *wbar, const double *bx, double *gax, const int *tuples, int t0, const int *xlist,
                const int *xoff, const int *ent, int nx, int no, int nup, int kp, void *stream);
size_t ccx_hip_gbx_part_doubles(int nup, int kp, int nent, int nsplit);
int ccx_hip_gbx(const double *wbar, const double *axr, const double *axc, double *gbx, double *part,
                const int *tuples, int t0, const int *plist, const int *poff, const int *ent, int npair, int nent,
                int nsplit, int no, int nup, int kp, void *stream);
int ccx_hip_dvoe_reduce(const double *slots, double *dvoep, const int *plist, const int *poff, const int *ent,
                        int np, int no, int nup, void *stream);
int ccx_hip_unpack_grad(const double *gax, const double *gbx, const double *dvoep, const double *v1, double *dvvvo,
                        double *dt2, double *dvm, double *dvoe, double *dt1, int no, int nu, int nup, int kp,
                        void *stream);
int ccx_hip_zero(void *p, size_t bytes, void *stream);

int ccx_hip_event_create(void **ev);
int ccx_hip_event_destroy(void *ev);
int ccx_hip_event_record(void *ev, void *stream);
int ccx_hip_event_ms(void *start, void *stop, float *ms);

#endif
