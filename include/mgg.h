/*
 * mgg.h - C ABI of libMultiGPUGridder.so, the MI355X-native Fourier-domain
 * projection gridder.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  Every
 * function returns MGG_OK (0) on success or a non-zero MGG_ERR_* code; the
 * message for the calling thread's last failure is mgg_last_error().
 *
 * Array layout (every array): column-major with x fastest, exactly as the
 * reference's Matlab callers hand them over:
 *   volume       vol[x + y*N + z*N*N]                 N^3 float
 *   CAS volume   cas[x + y*V + z*V*V]                 V^3 float, V = M + 2*pad
 *   images       img[a*N*N + x + y*N]                 P x N^2 float
 *   CAS images   cimg[a*M*M + x + y*M]                P x M^2 float, M = N*interp
 *   directions   axes[9*a + 0..8] = [u; v; n]         9 floats per direction:
 *                image x-axis, image y-axis, projection normal (volume frame).
 *
 * Four layers:
 *   1. the gridder object (mgg_create ... mgg_reconstruct_volume): the
 *      reference MultiGPUGridder / Matlab class surface.  Host arrays are
 *      borrowed (never copied at set time, never freed), results are written
 *      in place into the caller's arrays, and calls are synchronous;
 *   2. device-pointer operators (mgg_op_*): the same kernels on caller-owned
 *      device memory and a caller stream, no host staging;
 *   3. the legacy symbols the reference's ctypes wrapper binds
 *      (Python/MultiGPUGridder_Python.py:11-48);
 *   4. a request server (mgg_server_*): resident volumes and continuous
 *      batching of forward / back projection requests.
 */
#ifndef MGG_H
#define MGG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MGG_OK 0
#define MGG_ERR_INVALID_ARGUMENT 1 /* bad size, null pointer, bad option */
#define MGG_ERR_NOT_SET 2          /* a required array was never set */
#define MGG_ERR_DEVICE 3           /* no such device / device unusable */
#define MGG_ERR_OUT_OF_MEMORY 4    /* device or pinned allocation failed */
#define MGG_ERR_HIP 5              /* a HIP runtime call failed */
#define MGG_ERR_FFT 6              /* rocFFT plan or execution failed */
#define MGG_ERR_UNSUPPORTED 7      /* valid request this build does not do */
#define MGG_ERR_TIMEOUT 8          /* mgg_server_wait: request not finished in time */
#define MGG_ERR_BUSY 9             /* mgg_server_submit: request queue full */

/* Storage types a device CAS volume can be gathered from (mgg_op_forward_project). */
#define MGG_DTYPE_F32 0
#define MGG_DTYPE_BF16 1
#define MGG_DTYPE_F16 2

typedef struct mgg_gridder mgg_gridder;

/* Message of the calling thread's most recent failure ("" if none). */
const char *mgg_last_error(void);
/* Library version string. */
const char *mgg_version(void);
/* Number of visible GPUs (0 if the runtime finds none). */
int mgg_device_count(int *count);

/* ======================================================================== */
/* 1. Gridder object                                                         */
/* ======================================================================== */

/* Replaces MultiGPUGridder(VolumeSize, numCoordAxes, interpFactor,
 * extraPadding, Num_GPUs, GPU_Devices, RunFFTOnDevice, verbose)
 * (Gridders/MultiGPUGridder.h:36; Mex/mexCreateGridder.cpp:12-27).
 * Unlike the reference it does not reset the device, so several gridders can
 * share one GPU.  gpu_devices may be NULL when num_gpus is 1 (device 0). */
int mgg_create(int volume_size, int num_coord_axes, float interp_factor, int extra_padding,
               int num_gpus, const int *gpu_devices, int run_fft_on_device, int verbose,
               mgg_gridder **out);
/* Replaces delete / mexDeleteGridder (Mex/mexDeleteGridder.cpp:6). Frees only
 * this gridder's device memory. NULL is accepted. */
int mgg_destroy(mgg_gridder *g);

/* Setters.  Each replaces the mexSetVariables command of the same name
 * (Mex/mexSetVariables.cpp:27-97) and the AbstractGridder setter behind it
 * (Gridders/AbstractGridder.h:53-92).  Arrays are borrowed: they must stay
 * valid until re-set or mgg_destroy.  dims follow the Matlab size vector. */
int mgg_set_volume(mgg_gridder *g, float *volume, const int *dims /* [3] = N,N,N */);
int mgg_set_coord_axes(mgg_gridder *g, float *coord_axes, const int *dims /* [2] = 9P,1 */);
int mgg_set_images(mgg_gridder *g, float *images, const int *dims /* [3] = N,N,P */);
int mgg_set_cas_images(mgg_gridder *g, float *cas_images, const int *dims /* [3] = M,M,P */);
int mgg_set_cas_volume(mgg_gridder *g, float *cas_volume, const int *dims /* [3] = V,V,V */);
int mgg_set_kb_table(mgg_gridder *g, float *table, const int *dims /* [1] = K */);
int mgg_set_mask_radius(mgg_gridder *g, float mask_radius);
int mgg_set_num_axes(mgg_gridder *g, int num_axes);
int mgg_set_max_axes_to_allocate(mgg_gridder *g, int max_axes);
int mgg_set_num_streams_fp(mgg_gridder *g, int n);
int mgg_set_num_streams_bp(mgg_gridder *g, int n);
/* 'SetGPUs' (the reference silently ignores it, MultiGPUGridder_Matlab_Class.m:180).
 * Only valid before the first projection. */
int mgg_set_gpus(mgg_gridder *g, const int *gpu_devices, int num_gpus);
/* Kaiser-Bessel half width (AbstractGridder.cpp:10; default 2). */
int mgg_set_ker_hwidth(mgg_gridder *g, float hwidth);

/* Getters. */
int mgg_get_sizes(const mgg_gridder *g, int *volume_size, int *cas_image_size,
                  int *cas_volume_size, int *num_axes);
float mgg_get_mask_radius(const mgg_gridder *g);

/* Replaces MultiGPUGridder::ForwardProject (Gridders/MultiGPUGridder.cpp:98,
 * gpuGridder.cpp:180-305).  Gathers num_axes central slices from the CAS
 * volume into CAS images and, when run_fft_on_device is set, turns them into
 * images written into the caller's images array.  CAS images are written to
 * the caller's CAS-images array when one is set and run_fft_on_device is 0 or
 * verbose.  The CAS volume is the one last produced by
 * mgg_volume_to_cas_volume, else the caller's CAS-volume array (uploaded on
 * every call, as the reference does). */
int mgg_forward_project(mgg_gridder *g);
/* Replaces MultiGPUGridder::BackProject (MultiGPUGridder.cpp:186,
 * gpuGridder.cpp:307-430).  Zeroes the device CAS volume, then back projects
 * the images (run_fft_on_device = 1) or the CAS images (0) of num_axes
 * directions into it.  Per-stream partial volumes are summed in FP64. */
int mgg_back_project(mgg_gridder *g);
/* Replaces MultiGPUGridder::CASVolumeToVolume (MultiGPUGridder.cpp:271-304):
 * copies the device CAS volume into the caller's CAS-volume array. */
int mgg_cas_volume_to_host(mgg_gridder *g);
/* GPU form of utils/CASFromVol_Gridder.m:26-60 (run on the host by the
 * reference's forwardProject, MultiGPUGridder_Matlab_Class.m:226): caller's
 * volume -> device CAS volume (stays resident for later forward projections);
 * also written to the caller's CAS-volume array if one is set. */
int mgg_volume_to_cas_volume(mgg_gridder *g);
/* GPU form of utils/volFromCAS_Gridder.m:27-56 (the reference's getVol,
 * MultiGPUGridder_Matlab_Class.m:314-325): device CAS volume -> caller's
 * volume array. */
int mgg_get_volume(mgg_gridder *g);
/* GPU form of MultiGPUGridder_Matlab_Class.m:284-312 (reconstructVol; the
 * reference's mexMultiGPUReconstructVolume is an empty stub): back projects
 * the images and, in the same pass, the plane density (all-ones CAS images),
 * divides CAS(BP) / (CAS(density) + 1e-6) in FP64 and converts to the
 * caller's volume array. */
int mgg_reconstruct_volume(mgg_gridder *g);
/* Replaces resetVolume (MultiGPUGridder_Matlab_Class.m:277-282,
 * AbstractGridder.cpp:189-192): zeroes the caller's volume and CAS volume
 * arrays (when set) and the device CAS volume. */
int mgg_reset_volume(mgg_gridder *g);

/* ======================================================================== */
/* 2. Device-pointer operators                                               */
/* ======================================================================== */
/* All pointers are device pointers on the current device; stream is a
 * hipStream_t (NULL = the null stream).  Calls are asynchronous on stream.
 * ker is the device copy of the Kaiser-Bessel table (ker_size floats,
 * half width ker_hwidth).  Threads may call operators concurrently on
 * different streams; calls on one stream must come from one thread at a time
 * (the image and volume chains keep per-stream FFT plans and work memory). */

/* Forward projection (Projection/gpuForwardProject.cu:4-101): writes n_axes
 * CAS images of m x m (0 outside the mask disc).  cas_volume holds v^3
 * elements of cas_dtype (MGG_DTYPE_*). */
int mgg_op_forward_project(const void *cas_volume, int cas_dtype, int v, const float *axes,
                           int n_axes, float *cas_images, int m, float mask_radius,
                           const float *ker, int ker_size, float ker_hwidth, void *stream);
/* Writes *check (one unsigned in device memory) = 1 when ker holds 501 finite
 * entries whose first and last are 0 (as the reference's table,
 * getKernelFiltTable.m) and all v^3 elements of cas_volume are finite, else
 * 0; asynchronous on stream.  No reference counterpart: it licenses the
 * shortcut of mgg_op_forward_project_checked. */
int mgg_op_check_volume(const void *cas_volume, int cas_dtype, int v, const float *ker, int ker_size,
                        unsigned *check, void *stream);
/* mgg_op_forward_project with the flag mgg_op_check_volume wrote for this
 * cas_volume and ker (same stream or ordered after it, and after the volume's
 * last change; NULL = no flag).  Where the flag reads 1 each axis of a stencil
 * has one end tap whose weight is table entry 0 or 500 (both 0), and the
 * gather skips it: results equal mgg_op_forward_project's, the voxel boxes
 * staged per pixel tile are about 40% smaller.  Where it reads 0 this is
 * mgg_op_forward_project.  No reference counterpart. */
int mgg_op_forward_project_checked(const void *cas_volume, int cas_dtype, int v, const float *axes,
                                   int n_axes, float *cas_images, int m, float mask_radius,
                                   const float *ker, int ker_size, float ker_hwidth,
                                   const unsigned *check, void *stream);
/* Back projection with the reference's semantics (Projection/gpuBackProject.cu:
 * 4-113), accumulated into cas_volume (v^3 FP32) and, if density is non-NULL,
 * the plane density into density (v^3).  Owner-computes: launches that touch
 * the same cas_volume must be ordered on one stream. */
int mgg_op_back_project(float *cas_volume, float *density, int v, const float *axes,
                        int n_axes, const float *cas_images, int m, float mask_radius,
                        const float *ker, int ker_size, float ker_hwidth, void *stream);
/* Exact transpose of mgg_op_forward_project, accumulated into cas_volume
 * (no reference counterpart: the adjoint the autograd backward needs). */
int mgg_op_back_project_adjoint(float *cas_volume, int v, const float *axes, int n_axes,
                                const float *cas_images, int m, float mask_radius,
                                const float *ker, int ker_size, float ker_hwidth, void *stream);
/* mgg_op_check_volume for n_images CAS images of m x m (FP32): *check = 1
 * when ker holds 501 finite entries with zero ends and every image value is
 * finite, else 0 (cas_images may be NULL when n_images is 0).  No reference
 * counterpart. */
int mgg_op_check_images(const float *cas_images, int n_images, int m, const float *ker, int ker_size,
                        unsigned *check, void *stream);
/* mgg_op_back_project_adjoint with mgg_op_check_images' flag for these
 * cas_images and ker (NULL = no flag): where it reads 1 the scatter skips each
 * stencil axis' end tap of weight 0, as mgg_op_forward_project_checked does
 * (same sums, smaller boxes).  No reference counterpart. */
int mgg_op_back_project_adjoint_checked(float *cas_volume, int v, const float *axes, int n_axes,
                                        const float *cas_images, int m, float mask_radius,
                                        const float *ker, int ker_size, float ker_hwidth,
                                        const unsigned *check, void *stream);
/* Gridders/gpuGridder.cpp:432-518: n CAS images (m x m) -> images (nn x nn).
 * Uses a batched rocFFT C2R; cas_images is clobbered (used as scratch). */
int mgg_op_cas_images_to_images(float *cas_images, float *images, int n, int m, int nn,
                                void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mgg_op_cas_images_to_images without any allocation and
 * without clobbering cas_images.  Only the fused chains' sizes are accepted
 * (m = 128, 192, 256, 384 or 512, nn <= m, m - nn even; others return
 * MGG_ERR_UNSUPPORTED); the intermediate lives in a fixed 32 MiB device array
 * per GPU, held from library load to process end, and n is processed in chunks
 * that fit it.  Calls that overlap on one GPU share that array and race:
 * order them on one stream. */
int mgg_op_cas_images_to_images_noalloc(const float *cas_images, float *images, int n, int m, int nn,
                                        void *stream);
/* Gridders/gpuGridder.cpp:520-596: n images (nn x nn) -> CAS images (m x m). */
int mgg_op_images_to_cas_images(const float *images, float *cas_images, int n, int m, int nn,
                                void *stream);
/* utils/CASFromVol_Gridder.m: volume (nn^3) -> CAS volume (v^3, v = m+2*pad). */
int mgg_op_volume_to_cas_volume(const float *volume, int nn, float *cas_volume, int m, int pad,
                                float ker_hwidth, void *stream);
/* utils/volFromCAS_Gridder.m: CAS volume (v^3) -> volume (nn^3).
 * zero_nyquist = 1 is the reference's FromCAS; 0 keeps the index-0 planes
 * (the exact transpose of mgg_op_volume_to_cas_volume, up to m^3). */
int mgg_op_cas_volume_to_volume(const float *cas_volume, int m, int pad, float *volume, int nn,
                                float ker_hwidth, int zero_nyquist, void *stream);
/* Transposes for differentiation (no counterpart in the reference, which has
 * no gradients).  The full spatial projector is
 *   images = CASImgsToImgs(Gather(CASFromVol(volume)))
 * and its transpose is the composition of these in reverse order with
 * mgg_op_back_project_adjoint:
 *   cas_images_to_images_transpose: cas = (1/m^2) ImgsToCAS(images) with the
 *     index-0 row and column of each CAS image set to 0;
 *   volume_to_cas_volume_transpose: m^3 CASToVol(cas) keeping the index-0
 *     planes (= mgg_op_cas_volume_to_volume(zero_nyquist = 0) times m^3). */
int mgg_op_cas_images_to_images_transpose(const float *images, float *cas_images, int n, int m, int nn,
                                          void *stream);
int mgg_op_volume_to_cas_volume_transpose(const float *cas_volume, int m, int pad, float *volume, int nn,
                                          float ker_hwidth, void *stream);
/* out[i] = num[i] / (den[i] + eps), evaluated in FP64
 * (MultiGPUGridder_Matlab_Class.m:308). out may alias num. */
int mgg_op_divide_density(float *out, const float *num, const float *den, size_t n, double eps,
                           void *stream);
/* out[i] (+)= sum_p parts[p*stride + i] in FP64 (multi-stream / multi-device
 * partial-volume reduction, replaces MultiGPUGridder.cpp:376-411). */
int mgg_op_reduce_partials(float *out, const float *parts, int n_parts, size_t n, size_t stride,
                           int accumulate, void *stream);
/* FP32 CAS volume -> BF16 / FP16 storage copy (round to nearest even). */
int mgg_op_convert_volume(void *dst, int dtype, const float *src, size_t n, void *stream);

/* Standalone filters: the element semantics of the reference's Filters kernels. */
int mgg_filter_cas_to_complex(const float *in, float *out_complex, int size, int n_slices,
                              void *stream); /* CASToComplexFilter.cu:3-81; n_slices<=0: 3-D */
int mgg_filter_complex_to_cas(const float *in_complex, float *out, int size, int n_slices,
                              void *stream); /* ComplexToCASFilter.cu:3-26 */
int mgg_filter_complex_to_real(const float *in_complex, float *out, int size, int n_slices,
                               void *stream); /* ComplexToRealFilter.cu:3-25 */
int mgg_filter_real_to_complex(const float *in, float *out_complex, int size, int n_slices,
                               void *stream); /* RealToComplexFilter.cu:3-26 */
int mgg_filter_fftshift2d(void *data, int elem_bytes, int size, int n_slices,
                          void *stream); /* FFTShift2DFilter.cu:3-62, in place */
int mgg_filter_fftshift3d(void *data, int elem_bytes, int size,
                          void *stream); /* FFTShift3DFilter.cu:3-98, in place */
int mgg_filter_pad(const float *in, float *out, int in_size, int out_size, int n_slices,
                   void *stream); /* PadVolumeFilter.cu:4-36 */
int mgg_filter_crop(const void *in, void *out, int elem_bytes, int in_size, int out_size,
                    int n_slices, void *stream); /* CropVolumeFilter.cu:3-39 */
int mgg_filter_divide_scalar(float *data, float scalar, size_t n,
                             void *stream); /* DivideScalarFilter.cu:3-25 */
int mgg_filter_divide_volume(float *a, const float *b, size_t n,
                             void *stream); /* DivideVolumeFilter.cu:3-33 */
int mgg_filter_multiply_volume(float *a, const float *b, size_t n,
                               void *stream); /* MultiplyVolumeFilter.cu:29-53 */
int mgg_filter_multiply_complex_volume(float *a_complex, const float *b, size_t n,
                                       void *stream); /* MultiplyVolumeFilter.cu:3-27 */
int mgg_filter_add_volume(float *a, const float *b, size_t n,
                          void *stream); /* AddVolumeFilter.cu:3-27 */

/* Host math the reference computes in Matlab. */
/* Matlab/getKernelFiltTable.m:1-22 (+ gridLibKaiser.m, kernelAlpha.m). */
int mgg_kb_table(float *out, int size, int hwidth);
/* Matlab/getPreComp.m:1-28: out[m]. */
int mgg_precomp(float *out, int m, int hwidth);
/* Matlab/utils/create_uniform_axes.m:16-28 with get_axes_from_n_rot.m:7-37:
 * out[9 * ntheta * nphi]. */
int mgg_uniform_axes(float *out, int ntheta, int nphi);

/* ======================================================================== */
/* 3. Legacy ctypes symbols (Python/MultiGPUGridder_Python.py:11-48)         */
/* ======================================================================== */
/* Thin shims over layer 1 with the reference wrapper's names and argument
 * meaning.  They return the handle (or NULL on failure) where the wrapper
 * declares a void* result. */
void *Gridder_new(void);
void *SetNumberGPUs(void *h, int n);
void *SetNumberStreams(void *h, int n);
void *SetVolume(void *h, float *volume, int *dims);
float *GetVolume(void *h);
void *ResetVolume(void *h);
void *SetImages(void *h, float *images);
float *GetImages(void *h);
void *SetAxes(void *h, float *axes, int *dims);
void *SetImgSize(void *h, int *dims);
void *SetMaskRadius(void *h, float *radius);
void *Forward_Project(void *h);
void *Back_Project(void *h);

/* ======================================================================== */
/* 4. Request server                                                        */
/* ======================================================================== */
/* No counterpart in the reference, whose callers run one synchronous
 * forwardProject / backProject per request (MultiGPUGridder_Matlab_Class.m:198,
 * :245).  The server keeps CAS volumes resident in HBM and runs requests of
 * mixed N and direction counts:
 *   - admission: each batch's device memory comes from one arena of
 *     budget_bytes; a batch starts when a pipeline is free and it fits;
 *   - batching: forward requests on one volume share launches of up to
 *     max_batch_directions directions; a back request owns its accumulator;
 *   - overlap: num_pipelines batches in flight, one stream each.
 * Results equal the plain path's for the same request: a forward request
 * returns forwardProject's images of its directions (n_axes x N x N); a back
 * request returns getVol of the back projection of its images into a zeroed
 * volume (N^3).  A result does not depend on the other requests in flight.
 * Array arguments may be pageable host memory, pinned host memory
 * (mgg_host_alloc; copies then run asynchronously) or device memory; they
 * must stay valid until mgg_server_wait returns for the request.
 * Thread safe: any thread may submit and wait. */
typedef struct mgg_server mgg_server;

#define MGG_REQ_FORWARD 0
#define MGG_REQ_BACK 1

typedef struct mgg_server_config {
    int device;               /* HIP device ordinal */
    size_t budget_bytes;      /* device memory for volumes + in-flight batches
                                 (0: half the free memory, at most 32 GiB) */
    int max_batch_directions; /* directions per launch (0: 8192) */
    int num_pipelines;        /* batches in flight (0: 4) */
    int max_queued;           /* waiting requests before submit says BUSY (0: no limit) */
} mgg_server_config;

typedef struct mgg_request {
    int kind;          /* MGG_REQ_FORWARD or MGG_REQ_BACK */
    int volume_id;     /* from mgg_server_add_volume */
    int n_axes;        /* directions */
    const float *axes; /* 9 * n_axes */
    float *images;     /* forward: output n_axes x N^2; back: input */
    float *volume;     /* back: output N^3; forward: unused (NULL) */
} mgg_request;

typedef struct mgg_request_info {
    int status;      /* MGG_OK or the request's failure code */
    double t_submit; /* seconds on the server clock (mgg_server_clock) */
    double t_start;  /* first batch of the request started */
    double t_done;   /* result written to the caller's arrays */
} mgg_request_info;

typedef struct mgg_server_stats {
    unsigned long long submitted, completed, failed;
    unsigned long long batches;            /* launches of the projection kernels */
    unsigned long long batched_directions; /* directions over all batches */
    unsigned long long merged_batches;     /* forward batches holding several requests */
    size_t budget_bytes, resident_bytes, in_use_bytes, peak_bytes;
    int queued, in_flight, volumes;
} mgg_server_stats;

/* cfg may be NULL (all defaults, device 0). */
int mgg_server_create(const mgg_server_config *cfg, mgg_server **out);
/* Finishes every submitted request, then frees the server. */
int mgg_server_destroy(mgg_server *s);
/* volume: N^3 host floats (x fastest).  The CAS volume (CASFromVol_Gridder.m)
 * is built on the device and stays resident; MGG_ERR_OUT_OF_MEMORY if the
 * budget cannot hold it. */
int mgg_server_add_volume(mgg_server *s, const float *volume, int volume_size, float interp_factor,
                          int extra_padding, int *volume_id);
/* Waits until no request uses the volume, then frees it. */
int mgg_server_remove_volume(mgg_server *s, int volume_id);
/* Queues a request; *ticket identifies it for mgg_server_wait.  Fails at once
 * (nothing queued) on bad arguments, an unknown volume, a request larger than
 * the budget (MGG_ERR_OUT_OF_MEMORY) or a full queue (MGG_ERR_BUSY). */
int mgg_server_submit(mgg_server *s, const mgg_request *req, unsigned long long *ticket);
/* Waits for a request (timeout_s < 0: no limit) and returns its status; the
 * ticket is consumed unless the wait times out (MGG_ERR_TIMEOUT). */
int mgg_server_wait(mgg_server *s, unsigned long long ticket, double timeout_s, mgg_request_info *info);
int mgg_server_get_stats(mgg_server *s, mgg_server_stats *stats);
/* hold != 0: submitted requests queue up and no new batch starts (running
 * batches finish); hold = 0 lets the scheduler batch everything queued.  A
 * caller uses it to hand over a group of requests that should share launches.
 * mgg_server_destroy releases a hold. */
int mgg_server_hold(mgg_server *s, int hold);
/* Seconds since the server was created (the clock of mgg_request_info). */
double mgg_server_clock(mgg_server *s);
/* Pinned host memory, so request transfers run asynchronously. */
int mgg_host_alloc(size_t bytes, void **ptr);
int mgg_host_free(void *ptr);

#ifdef __cplusplus
}
#endif

#endif /* MGG_H */
