/*
 * mkgpu.h -- C ABI of libmkgpu.so, the MI355X (gfx950) kernels behind the
 * minkasi mapmaking inner loop.
 *
 * Conventions (all entry points):
 *   - Pointers named dev_* or documented as device memory are HIP device
 *     pointers; every other pointer argument is host memory.
 *   - Sizes are int64_t (the reference computes ndet*ndata in int, which
 *     overflows past 2^31 samples; minkasi/lib/minkasi.c:52).
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *     Calls are asynchronous on that stream unless stated otherwise.
 *   - The return value is a status (MK_OK == 0).  On failure,
 *     mk_last_error() returns a message for the calling thread.  This is a
 *     deliberate divergence from the reference, whose C functions return void
 *     and check nothing.
 *   - No entry point takes or returns a torch type.
 */
#ifndef MKGPU_H
#define MKGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MK_OK 0
#define MK_ERR_ARG 1     /* invalid argument (null pointer, bad size) */
#define MK_ERR_HIP 2     /* HIP runtime error */
#define MK_ERR_RANGE 3   /* index out of range (pixel outside the map) */
#define MK_ERR_NUMERIC 4 /* non-finite value or zero curvature */
#define MK_ERR_ALLOC 5   /* device allocation failed */

/* Message for the last failing call on this thread. */
const char *mk_last_error(void);
/* Library ABI version (increments when a signature changes). */
int mk_version(void);
/* Number of bytes of workspace the reductions below need (device memory,
 * 256-byte aligned; contents may be garbage on entry, no clearing needed). */
int64_t mk_reduce_workspace_bytes(void);

/* ------------------------------------------------------------------ */
/* Pointing -> pixel.
 * Replaces SkyMap.pix_from_radec (minkasi/maps/skymap.py:457-490), which
 * calls astropy wcs_world2pix for a CAR WCS built by get_wcs
 * (minkasi/maps/utils.py:47-66), subtracts 1, rounds half-to-even and
 * stores int32(x*ny + y).
 * wcs10 (host, 10 doubles): euler[6] = {alpha_p, 90-delta_p, phi_p,
 *   cos(90-delta_p), sin(90-delta_p), car_scale}, cdelt[2], crpix[2].
 * ra, dec: device, radians, nsamp doubles each.  ipix: device int32 out.
 * dev_nbad (device int64, may be NULL): number of samples whose pixel falls
 * outside [0,nx)x[0,ny) is ADDED to it (the reference has no bounds check;
 * those samples get the index the formula gives, as in the reference). */
int mk_pix_from_radec_car(const double *ra, const double *dec, int64_t nsamp,
                          const double *wcs10, int32_t nx, int32_t ny,
                          int32_t *ipix, int64_t *dev_nbad, void *stream);
/* Same rotation, raw 1-based pixel coordinates (wcs_world2pix(..., 1)),
 * for the near-tie audit.  px, py: device doubles. */
int mk_world2pix_car(const double *ra, const double *dec, int64_t nsamp,
                     const double *wcs10, double *px, double *py, void *stream);
/* Pointing of a separable raster, ra[d,t] = ra_b[t] + dra[d] and
 * dec[d,t] = dec_b[t] + ddec[d], pixelized without materialising ra/dec
 * (ndet x ndata samples, row-major).  Same arithmetic per sample as
 * mk_pix_from_radec_car. */
int mk_pix_from_raster_car(const double *ra_b, const double *dec_b,
                           const double *dra, const double *ddec,
                           int64_t ndet, int64_t ndata, const double *wcs10,
                           int32_t nx, int32_t ny, int32_t *ipix,
                           int64_t *dev_nbad, void *stream);

/* ------------------------------------------------------------------ */
/* map2tod: dat[i] (+)= map[ipix[i]] for i < nsamp.
 * Replaces map2tod_omp / map2tod_simple (minkasi/lib/minkasi.c:116-144)
 * bound by minkasi/lib/minkasi.py (argtypes) and called from
 * minkasi/mapmaking/map2tod.py:19-54.  nsamp = ndet*ndata of one TOD or of
 * a whole batch of concatenated TODs.  Bit-exact (a gather).  Any
 * element-aligned ipix / dat: 16-byte aligned ipix with dat aligned to 4
 * elements take 4-wide loads and stores, others one sample per thread. */
int mk_map2tod_f32(float *dat, const float *map, int64_t nsamp,
                   const int32_t *ipix, int do_add, void *stream);
int mk_map2tod_f64(double *dat, const double *map, int64_t nsamp,
                   const int32_t *ipix, int do_add, void *stream);

/* Tiled map layout for the map2tod gather (an MI355X layout; the
 * reference's map2tod_omp, minkasi.c:131-144, gathers from the row-major
 * map).  Pixel (x, y) of an nx x ny map is stored at
 *   ((x / tx) * ceil(ny / ty) + y / ty) * tx * ty + (x % tx) * ty + y % ty,
 * so a raster track touches a few 128-byte lines per wave instead of one
 * per map row.  mk_map_tiled_size gives the tiled array's length (-1 on bad
 * sizes); tx * ty <= 1024.  mk_pix_tile_remap rewrites row-major indices
 * (x * ny + y) into tiled ones; indices outside the map are counted into
 * *dev_nbad (device int64, may be NULL) and written as 0.
 * mk_map_to_tiles_* copies a row-major map into the tiled array (padding
 * zero).  mk_map2tod_* on (tiled map, remapped indices) equals mk_map2tod_*
 * on (map, indices) bit for bit. */
/* Pixel-index validation (no reference counterpart: the reference's
 * map2tod_omp / tod2map_* index the map unchecked, minkasi/lib/minkasi.c:
 * 50-144).  Adds to *dev_nbad (device int64, caller-zeroed) the number of
 * the nsamp indices outside [0, npix).  The Python layer runs it once per
 * saved pixelization before any map2tod / atomic tod2map reads it. */
int mk_pix_count_bad(const int32_t *ipix, int64_t nsamp, int64_t npix,
                     int64_t *dev_nbad, void *stream);

int64_t mk_map_tiled_size(int32_t nx, int32_t ny, int32_t tile_x,
                          int32_t tile_y);
int mk_pix_tile_remap(const int32_t *ipix, int64_t nsamp, int32_t nx,
                      int32_t ny, int32_t tile_x, int32_t tile_y,
                      int32_t *out, int64_t *dev_nbad, void *stream);
int mk_map_to_tiles_f32(const float *map, int32_t nx, int32_t ny,
                        int32_t tile_x, int32_t tile_y, float *tiled,
                        void *stream);
int mk_map_to_tiles_f64(const double *map, int32_t nx, int32_t ny,
                        int32_t tile_x, int32_t tile_y, double *tiled,
                        void *stream);

/* Coded pointing for the tiled gather: map2tod_omp (minkasi.c:131-144)
 * reading about 1 byte of index per sample instead of 4.  The arena's
 * samples are cut into blocks of 256; a block keeps its first sample's pixel
 * (x0, y0) and per sample the offset from it, as int8 dx with dy = 0
 * (type 1, 1 B), int8 (dx, dy) (type 2, 2 B) or packed (x << 16 | y)
 * (type 3, 4 B).  Map sides must be below 32768.
 *  mk_pix_code_blocks: number of blocks, ceil(nsamp / 256).
 *  mk_pix_code_size: per block, payload size in 256-byte units (1, 2 or 4)
 *      into units (device int32, nblocks); indices outside [0, nx*ny) are
 *      counted into *dev_nbad (device int64, caller-zeroed).
 *  mk_pix_code_write: with off_units the exclusive prefix sum of units
 *      (device int32, total below 2^30), writes the headers (hdr: device,
 *      8 bytes per block) and the payload (256-byte aligned, 256 * total
 *      units bytes).
 *  mk_map2tod_coded_*: dat (+)= tiled[index(x, y)] over the arena, with
 *      tiled the map copy of mk_map_to_tiles_* for tile_x x tile_y tiles
 *      (powers of two); equal bit for bit to mk_map2tod_* on the map and
 *      the original indices.  dat 4-element aligned. */
int64_t mk_pix_code_blocks(int64_t nsamp);
int mk_pix_code_size(const int32_t *ipix, int64_t nsamp, int32_t nx, int32_t ny,
                     int32_t *units, int64_t *dev_nbad, void *stream);
int mk_pix_code_write(const int32_t *ipix, int64_t nsamp, int32_t ny,
                      const int32_t *off_units, const int32_t *units,
                      void *hdr, void *payload, void *stream);
int mk_map2tod_coded_f32(float *dat, const float *tiled, int64_t nsamp,
                         const void *hdr, const void *payload, int32_t ny,
                         int32_t tile_x, int32_t tile_y, int do_add,
                         void *stream);
int mk_map2tod_coded_f64(double *dat, const double *tiled, int64_t nsamp,
                         const void *hdr, const void *payload, int32_t ny,
                         int32_t tile_x, int32_t tile_y, int do_add,
                         void *stream);

/* tod2map with global atomics: map[ipix[i]] += dat[i].
 * Counterpart of tod2map_atomic (minkasi.c:59-66); order-nondeterministic
 * like the reference's tod2map_omp/tod2map_atomic.  Kept as the
 * non-plan variant; the deterministic path is the plan below. */
int mk_tod2map_atomic_f32(float *map, const float *dat, int64_t nsamp,
                          const int32_t *ipix, void *stream);
int mk_tod2map_atomic_f64(double *map, const double *dat, int64_t nsamp,
                          const int32_t *ipix, void *stream);

/* Deterministic tod2map.  Replaces tod2map_omp (minkasi.c:97-112, bound in
 * minkasi/lib/minkasi.py, called from minkasi/mapmaking/tod2map.py:97-132
 * via SkyMap.tod2map, minkasi/maps/skymap.py:569-609).
 * A plan is built once per pixelization: samples are sorted by map tile
 * (stable), every 64-sample group is ordered by pixel, and the work is cut
 * into items.  Executing the plan sums each pixel's samples in a fixed
 * order, so the map is bitwise reproducible.  tile_y must be a multiple of
 * 4 and tile_x * tile_y at most 1024; item_samples (samples of one tile per
 * workgroup) must be a positive multiple of 64.  The plan owns device memory
 * allocated at creation; exec allocates nothing. */
typedef struct mk_t2m_plan mk_t2m_plan;
int mk_t2m_plan_create(const int32_t *ipix, int64_t nsamp, int32_t nx,
                       int32_t ny, int32_t tile_x, int32_t tile_y,
                       int32_t item_samples, mk_t2m_plan **plan,
                       void *stream);
/* map (+)= tod2map(dat).  accumulate=0 overwrites the map (pixels with no
 * samples become 0), accumulate=1 adds (the reference's do_add=True). */
int mk_t2m_plan_exec_f32(const mk_t2m_plan *plan, float *map, const float *dat,
                         int accumulate, void *stream);
int mk_t2m_plan_exec_f64(const mk_t2m_plan *plan, double *map,
                         const double *dat, int accumulate, void *stream);
/* mk_t2m_plan_exec_f32 with the sums kept in FP64: FP32 samples, FP64
 * segmented scans, LDS tiles and partial tiles (FP32 partials when the plan
 * could not afford 8-byte ones), rounded once into the map, which is FP32
 * (map_f64 = 0) or FP64 (map_f64 = 1).  Same plan, same fixed order, so
 * bitwise repeatable.  The reference sums in FP64 (minkasi.c:97-112 on
 * double maps); FP32 sums of a pixel's many cancelling samples were the
 * largest error of the FP32 operator at the PCG solution. */
int mk_t2m_plan_exec_f32_acc64(const mk_t2m_plan *plan, void *map,
                               int map_f64, const float *dat, int accumulate,
                               void *stream);
/* Deterministic polarised binning through the same plan: map (nx*ny x
 * ncomp, component-minor, PolMap's layout) (+)= the per-sample terms of mode
 * (MK_POL_*: QU -> d cos, d sin; IQU -> d, d cos, d sin; QU_PRECON ->
 * d cos^2, d sin^2, d sin cos; IQU_PRECON -> d, d cos, d sin, d cos^2,
 * d sin cos, d sin^2, with cos, sin of twogamma), one binning pass per
 * component in the plan's fixed order: bitwise repeatable, where the
 * reference's tod2map_{qu,iqu}[_precon]_simple (minkasi.c:178-258) sum
 * serially and mk_tod2polmap_atomic_* sums in arrival order.  FP32 samples
 * are summed in FP64. */
int mk_t2m_plan_exec_pol_f32(const mk_t2m_plan *plan, float *map,
                             const float *dat, const float *twogamma, int mode,
                             int accumulate, void *stream);
int mk_t2m_plan_exec_pol_f64(const mk_t2m_plan *plan, double *map,
                             const double *dat, const double *twogamma,
                             int mode, int accumulate, void *stream);
/* info[0..7] = nsamp, nitems, ntiles_touched, tile_x, tile_y, plan bytes,
 * partial-buffer bytes, max samples per item. */
int mk_t2m_plan_info(const mk_t2m_plan *plan, int64_t *info);
int mk_t2m_plan_destroy(mk_t2m_plan *plan);

/* ------------------------------------------------------------------ */
/* Map-space vector operations for PCG (minkasi/maps/skymap.py:428-440
 * axpy, :651-666 dot, :724-740 __mul__; minkasi/maps/mapset.py:89-208).
 * Maps are FP32; every reduction accumulates in FP64 with a fixed
 * two-stage order (bitwise reproducible).  `work` is a device workspace of
 * mk_reduce_workspace_bytes() bytes.  Device scalar block sc (FP64): */
#define MK_SC_ZR 0     /* r.z */
#define MK_SC_PAP 1    /* p.Ap */
#define MK_SC_ALPHA 2  /* zr / pAp */
#define MK_SC_BETA 3   /* zr_new / zr */
#define MK_SC_RR 4     /* r.r */
#define MK_SC_FLAG 5   /* nonzero: zero/negative curvature or non-finite */
#define MK_SC_DOT 6    /* result of mk_dot_* */
#define MK_SC_COUNT 8
/* sc[MK_SC_DOT] = sum a*b (FP64 accumulation). */
int mk_dot_f32(const float *a, const float *b, int64_t n, double *sc,
               void *work, void *stream);
int mk_dot_f64(const double *a, const double *b, int64_t n, double *sc,
               void *work, void *stream);
/* y += a*x with a a host scalar, evaluated in FP64 (SkyMap.axpy,
 * skymap.py:428-440; CutsCompact.axpy, cuts.py:445-457). */
int mk_axpy_f32(float *y, const float *x, double a, int64_t n, void *stream);
int mk_axpy_f64(double *y, const double *x, double a, int64_t n,
                void *stream);
/* out = a*b elementwise (SkyMap.__mul__ skymap.py:724-740, Jacobi
 * preconditioner, CutsCompact.__mul__ cuts.py:483-495). */
int mk_mul_f32(float *out, const float *a, const float *b, int64_t n,
               void *stream);
int mk_mul_f64(double *out, const double *a, const double *b, int64_t n,
               void *stream);
/* y += a*b elementwise (CutsCompact.apply_prior, cuts.py:459-481). */
int mk_addmul_f32(float *y, const float *a, const float *b, int64_t n,
                  void *stream);
int mk_addmul_f64(double *y, const double *a, const double *b, int64_t n,
                  void *stream);
/* PCG steps of minkasi/mapmaking/pcg.py:143-170 with device scalars:
 *  mk_pcg_alpha:  pAp = p.Ap ; alpha = zr/pAp  (flag if pAp <= 0 / NaN)
 *  mk_pcg_xr:     x += alpha p ; r -= alpha Ap ; rr = r.r
 *  mk_pcg_xrz_jacobi: as mk_pcg_xr, then z = w*r and zr_new = r.z,
 *                 beta = zr_new/zr, zr = zr_new (one fused pass)
 *  mk_pcg_beta:   zr_new = r.z ; beta = zr_new/zr ; zr = zr_new
 *  mk_pcg_p:      p = z + beta p
 *  mk_pcg_init:   zr = r.z ; rr = r.r ; p = z                          */
int mk_pcg_alpha(const float *p, const float *Ap, int64_t n, double *sc,
                 void *work, void *stream);
int mk_pcg_xr(float *x, float *r, const float *p, const float *Ap, int64_t n,
              double *sc, void *work, void *stream);
int mk_pcg_xrz_jacobi(float *x, float *r, float *z, const float *p,
                      const float *Ap, const float *w, int64_t n, double *sc,
                      void *work, void *stream);
int mk_pcg_beta(const float *r, const float *z, int64_t n, double *sc,
                void *work, void *stream);
int mk_pcg_p(float *p, const float *z, int64_t n, const double *sc,
             void *stream);
int mk_pcg_init(const float *r, const float *z, float *p, int64_t n,
                double *sc, void *work, void *stream);

/* Reliable updates for the true residual ||b - Ax||/||b|| (pcg.py:143-170
 * keeps x and r in FP64 numpy maps; here they are FP32 device maps):
 *  mk_pcg_xr_x64 / mk_pcg_xrz_jacobi_x64: as mk_pcg_xr / mk_pcg_xrz_jacobi,
 *      with x accumulated in the FP64 map x64 (x receives its FP32 rounding),
 *      so the iterate keeps the precision the FP32 map would lose.
 *  mk_pcg_set_r: r = fl32(b - ax) from the FP64 right-hand side b and an FP64
 *      product ax = A x64 (residual replacement), rr = r.r into the scalar
 *      block.  After it the recursive residual starts again from the true
 *      one, so its drift from the FP32 operator's rounding is reset. */
int mk_pcg_xr_x64(double *x64, float *x, float *r, const float *p,
                  const float *Ap, int64_t n, double *sc, void *work,
                  void *stream);
int mk_pcg_xrz_jacobi_x64(double *x64, float *x, float *r, float *z,
                          const float *p, const float *Ap, const float *w,
                          int64_t n, double *sc, void *work, void *stream);
int mk_pcg_set_r(float *r, const double *b, const double *ax, int64_t n,
                 double *sc, void *work, void *stream);


/* ------------------------------------------------------------------ */
/* DCT-I (FFTW REDFT00, unnormalised): rows of length n >= 2.
 * Replaces many_fft_r2r_1d / many_fftf_r2r_1d (minkasi/lib/mkfftw.c:174-254,
 * bound in minkasi/lib/mkfftw.py, called by fft_r2r in
 * minkasi/tools/fft.py:290-344; scipy fallback minkasi/tools/py_fft.py:239).
 * Bluestein chirp-z on a power-of-two four-step FFT; any n.  x, y: device,
 * nrow*n elements; y may equal x.  work: device workspace of at least
 * mk_dct1_workspace_bytes(nrow, n, sizeof element) bytes, 256-B aligned.
 * Synchronises the stream once (descriptor upload). */
int64_t mk_dct1_workspace_bytes(int64_t nrow, int64_t n, int elem_bytes);
int mk_dct1_f64(const double *x, double *y, int64_t nrow, int64_t n,
                void *work, int64_t work_bytes, void *stream);
/* mk_dct1_f64 without a caller workspace: no counterpart in the reference
 * (its FFTW plans allocate, minkasi/lib/mkfftw.c:174-212).  For callers that
 * cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc): the workspace is a static 64 MiB device array of the library,
 * one per GPU, so nrow and n must satisfy
 * mk_dct1_workspace_bytes(nrow, n, 8) <= 64 MiB (else MK_ERR_ARG with the
 * sizes; e.g. 30 rows at n = 65536).  Allocates nothing; synchronises the
 * stream once, as mk_dct1_f64.  Calls that overlap on one GPU share that
 * array and race: serialise them (one stream, or the caller's lock). */
int mk_dct1_f64_noalloc(const double *x, double *y, int64_t nrow, int64_t n,
                        void *stream);
int mk_dct1_f32(const float *x, float *y, int64_t nrow, int64_t n,
                void *work, int64_t work_bytes, void *stream);

/* Noise filter y = C W C x per detector mode, C = DCT-I, W = mywt: the
 * DCT-I / weight / DCT-I core of NoiseSmoothedSVD.apply_noise
 * (minkasi/mapmaking/noise.py:705-708), batched over every mode of every
 * TOD.  Computed as one circular convolution per row through a real FFT of
 * L >= 3(n-1)+1 points (exact for any n): L = 2^11 P1 with P1 5-smooth in
 * [2, 256] on the FP32 path (by default the smallest such L), else
 * L = 2^q.
 * desc (host): 6 int64 per segment (TOD): rows, n, in_off, in_stride,
 *   out_off, out_stride (element offsets into the exec in/out arrays).
 * elem_bytes: 4 (FP32 transform) or 8 (FP64 transform).
 * The plan owns its spectra (allocated by mk_filt_plan_set_weights; a plan
 * whose segments all borrow spectra holds none) and a work buffer of
 * rows*P complex values, allocated by the first call that needs it (so
 * capture a graph only after one eager exec) and freed by
 * mk_filt_plan_release_work, or supplied by the caller
 * (mk_filt_plan_set_work). */
typedef struct mk_filt_plan mk_filt_plan;
int mk_filt_plan_create(int nseg, const int64_t *desc, int elem_bytes,
                        mk_filt_plan **plan);
/* As mk_filt_plan_create, with the column length P1 of every segment chosen
 * by the caller (an autotuner; the FFTW-wisdom counterpart of
 * minkasi/tools/fft.py:347-384): p1[i] = 0 keeps the default, otherwise it
 * must be one of the FP32 hot lengths (mk_filt_hot_p1) with
 * 2^11 p1[i] >= the segment's minimum length.  Plans whose spectra are
 * borrowed must agree on P1. */
int mk_filt_plan_create_p1(int nseg, const int64_t *desc, int elem_bytes,
                           const int32_t *p1, mk_filt_plan **plan);
/* Writes up to cap hot column lengths (ascending) to out (may be NULL) and
 * returns how many there are. */
int mk_filt_hot_p1(int32_t *out, int cap);
/* wdct: device FP64, rows of C W (the DCT-I of each row of mywt), laid out
 * exactly like the exec input (in_off/in_stride of desc). */
int mk_filt_plan_set_weights(mk_filt_plan *plan, const double *wdct,
                             void *stream);
/* Segment `seg` of `plan` uses the spectrum of segment `src_seg` of `src`
 * (same rows, n and precision) instead of its own; `src` must outlive
 * `plan`.  Lets one batched plan reuse the per-TOD noise models' spectra. */
int mk_filt_plan_borrow_spectrum(mk_filt_plan *plan, int seg,
                                 const mk_filt_plan *src, int src_seg);
/* halve_edges=1 multiplies output samples 0 and n-1 by 1/2, which is the
 * reference's dd[:,0] *= .5; dd[:,-1] *= .5 (noise.py:710-711) moved before
 * the back-rotation it commutes with. */
int mk_filt_plan_set_halve(mk_filt_plan *plan, int halve_edges);
int mk_filt_plan_exec_f32(const mk_filt_plan *plan, const float *in,
                          float *out, void *stream);
/* Frees the work buffer (synchronises the stream first); a caller's buffer
 * (mk_filt_plan_set_work) is only detached. */
int mk_filt_plan_release_work(mk_filt_plan *plan, void *stream);
/* Bytes of work buffer the plan needs (rows * P complex values). */
int mk_filt_plan_work_bytes(const mk_filt_plan *plan, int64_t *bytes);
/* Makes the plan use the caller's device buffer (>= mk_filt_plan_work_bytes,
 * 16-byte aligned) as its work buffer instead of one of its own; NULL
 * detaches it.  Plans run one after another on one stream can share a
 * buffer, which bounds the work memory of a large batch: the batched
 * operator splits its TODs into groups that way.  No counterpart in the
 * reference, whose FFTs use per-call numpy/FFTW buffers
 * (minkasi/tools/fft.py:50-124). */
int mk_filt_plan_set_work(mk_filt_plan *plan, void *work, int64_t bytes,
                          void *stream);
int mk_filt_plan_exec_f64(const mk_filt_plan *plan, const double *in,
                          double *out, void *stream);
/* info[0..7] = nseg, K1/K3 items, K2 items, work bytes, spectrum bytes,
 * elem_bytes, LDS bytes K1/K3, LDS bytes K2. */
int mk_filt_plan_info(const mk_filt_plan *plan, int64_t *info);
/* geom[4*i..4*i+3] = P1, P2, columns per item, P of segment i. */
int mk_filt_plan_geometry(const mk_filt_plan *plan, int64_t *geom);
int mk_filt_plan_destroy(mk_filt_plan *plan);

/* ------------------------------------------------------------------ */
/* Detector rotation of apply_noise, dat_rot = v.dat and dd = vT.dat_rot
 * (minkasi/mapmaking/noise.py:704, :709; numpy BLAS in the reference), as a
 * split-BF16 MFMA GEMM with FP32 accumulation, batched over TODs.
 * A (v or vT) is stored as hi/mid/lo bf16 planes (hi = bf16(a),
 * mid = bf16(a-hi), lo = bf16(a-hi-mid)), row-major, padded to
 * M_pad = ceil(M/128)*128 rows and lda = ceil(K/32)*32 columns (zeros).
 * mk_split_bf16_padded_dims returns M_pad*lda and sets *M_pad, *lda. */
int64_t mk_split_bf16_padded_dims(int64_t M, int64_t K, int64_t *M_pad,
                                  int64_t *lda);
int mk_split_bf16_matrix(const double *a, int64_t M, int64_t K,
                         int64_t lda_src, void *hi, void *mid, void *lo,
                         void *stream);
/* The same padded layout as two fp16 planes of 2^aexp * a (hi = fp16(x),
 * lo = fp16(x - hi), x = 2^aexp * a); the caller picks aexp so that
 * |x| < 2^14 (|aexp| <= 512). */
int mk_split_f16_matrix(const double *a, int64_t M, int64_t K,
                        int64_t lda_src, int aexp, void *hi, void *lo,
                        void *stream);
/* Batched C = A.B.  desc: 11 int64 per segment: M, K, N, A_hi, A_mid,
 * A_lo (device addresses), lda, b_off, b_ld, c_off, c_ld (element offsets
 * into the exec-time FP32 base pointers).  mode 0: three bf16 products
 * (hh + hm + mh, ~2^-18 relative per term); mode 1: six bf16 products
 * (adds hl + lh + mm, ~FP32 accuracy); mode 2: three fp16 products
 * (hh + hl + lh of the fp16 pairs, ~FP32 accuracy), with fields 3, 4 the
 * planes of mk_split_f16_matrix and field 5 its aexp; each column of B is
 * scaled by a power of two that follows the column's largest magnitude
 * before it is split (accumulators are rescaled exactly when it changes),
 * and the scalings are undone exactly in the output.  B and C must not
 * overlap (all modes). */
typedef struct mk_rot_plan mk_rot_plan;
int mk_rot_plan_create(int nseg, const int64_t *desc, int mode,
                       mk_rot_plan **plan);
int mk_rot_plan_exec(const mk_rot_plan *plan, const float *b_base,
                     float *c_base, void *stream);
int mk_rot_plan_destroy(mk_rot_plan *plan);
/* Deflation of dominant detector modes ahead of the FP32 rotation (no
 * counterpart in the reference, which rotates in FP64: noise.py:704).  For
 * each segment (a TOD of ndet rows of n samples at element offset off) and
 * its leading k <= 8 eigenvectors U (ndet x k, the first k rows of v
 * transposed): exec stores c = U^T x (FP64 sums) in the plan and writes
 * out = x - U c (out may equal x); addback adds c to the first k rows of the
 * rotated y = v.out, which gives v.x exactly in exact arithmetic and keeps
 * the FP32 rotation free of the cancellation a strong common mode causes.
 * desc (host): 4 int64 per segment: ndet, n, off, k.  u_host: every
 * segment's U in order, FP64 row-major.  exec and addback must pair up (c
 * lives in the plan). */
typedef struct mk_defl_plan mk_defl_plan;
int mk_defl_plan_create(int nseg, const int64_t *desc, const double *u_host,
                        mk_defl_plan **plan);
int mk_defl_plan_exec(const mk_defl_plan *plan, const float *in, float *out,
                      void *stream);
int mk_defl_plan_addback(const mk_defl_plan *plan, float *y, void *stream);
int mk_defl_plan_destroy(mk_defl_plan *plan);
/* FP64 setup GEMMs (NoiseSmoothedSVD.__init__, noise.py:650 and :667):
 * C = A.B (A FP64; B FP32 if b_is_f32 else FP64), and the Gram matrix
 * G = D.D^T (M x M) that replaces the SVD's left singular vectors
 * (eigenvectors of D.D^T).  work: mk_gram_workspace_bytes(M, n) bytes. */
int mk_gemm_nn_f64(int64_t M, int64_t N, int64_t K, const double *A,
                   int64_t lda, const void *B, int b_is_f32, int64_t ldb,
                   double *C, int64_t ldc, void *stream);
int64_t mk_gram_workspace_bytes(int64_t M, int64_t n);
int mk_gram_f64(int64_t M, int64_t n, const void *D, int d_is_f32,
                int64_t ldd, double *G, void *work, int64_t work_bytes,
                void *stream);
/* C = A.B^T (M x N, FP64) for row-major A (M x n) and B (N x n) with a long
 * inner axis n, split over n-chunks and summed in chunk order: the
 * np.dot(tmp, uu.T) cross products of gapfill_eig
 * (minkasi/tods/processing.py:357-421).  work: mk_cross_workspace_bytes
 * bytes. */
int64_t mk_cross_workspace_bytes(int64_t M, int64_t N, int64_t n);
int mk_cross_f64(int64_t M, int64_t N, int64_t n, const double *A,
                 int64_t lda, const double *B, int64_t ldb, double *C,
                 void *work, int64_t work_bytes, void *stream);
/* Spectrum smoothing (minkasi/tools/smooth.py:49-78 smooth_many_vecs and
 * noise.py:674-678): out_k = sum_{t=t_lo..t_hi} g[|t|] P(k-t) over the even
 * periodic extension of each row (period 2(n-1)); P = x^2 if flags&1 else
 * x; flags&2 inverts the result and zeroes bin 0.  out32 (may be NULL)
 * receives an FP32 copy.  g: device FP64, max(|t_lo|, t_hi)+1 taps. */
int mk_smooth_rows_f64(const double *x, double *out, float *out32,
                       int64_t nrow, int64_t n, const double *g, int32_t t_lo,
                       int32_t t_hi, int32_t flags, void *stream);

/* x -= (u.x * inv_uu) u with u.x accumulated in FP64 into sc[MK_SC_DOT]:
 * removes the component of x along u (inv_uu = 1/u.u).  With u the
 * indicator of hit pixels it projects A.p off the null mode of A (the
 * constant on hit pixels); an exact no-op in exact arithmetic, it keeps FP32
 * PCG consistent.  No reference counterpart (minkasi/mapmaking/pcg.py:143-170
 * runs in FP64). */
int mk_project_out_f32(float *x, const float *u, int64_t n, double inv_uu,
                       double *sc, void *work, void *stream);

/* ---- Learned preconditioner (no reference counterpart; it plugs into the
 * reference's preconditioner hook `precon * r`, minkasi/mapmaking/pcg.py:
 * 110-111,157-158).  M = S (alpha I + K^T K) S, S = diag(sqrt(1/hits)),
 * K = chain of linear dilated 3x3 convolutions.  Planar FP32 activations
 * [B][C][H][W]; C is 1 or 16; weights w[Cout][Cin][9], tap d = 3*di + dj
 * at offset (dil*(di-1), dil*(dj-1)); zero padding. */
/* y (+)= conv(x, w) with dilation dil. */
int mk_conv3x3_f32(const float *x, int32_t B, int32_t Cin, int32_t H,
                   int32_t W, const float *w, int32_t Cout, int32_t dil,
                   float *y, int32_t accumulate, void *stream);
/* wt[c][o][8-d] = w[o][c][d]: weights of the adjoint layer. */
int mk_conv3x3_adjoint_weights(const float *w, int32_t Co, int32_t Ci,
                               float *wt, void *stream);
/* g[o][c][d] (+)= sum_{b,p} a[b,o,p] x[b,c,p+off_d]  (weight gradient). */
int64_t mk_conv3x3_wgrad_workspace_bytes(int32_t B, int32_t Co, int32_t Ci,
                                         int32_t H, int32_t W, int32_t dil);
int mk_conv3x3_wgrad_f32(const float *a, int32_t Co, const float *x,
                         int32_t Ci, int32_t B, int32_t H, int32_t W,
                         int32_t dil, float *g, int32_t accumulate,
                         void *work, int64_t work_bytes, void *stream);
/* Box restriction by f with stride f: y[b][I][J] = f^-2 * sum of the f x f
 * fine pixels of block (I, J) (zero beyond H, W); y is B x ceil(H/f) x
 * ceil(W/f).  prolong is its adjoint: x[b][i][j] (+)= f^-2 y[b][i/f][j/f]. */
int mk_restrict_box_f32(const float *x, int32_t B, int32_t H, int32_t W,
                        int32_t f, float *y, void *stream);
int mk_prolong_box_f32(const float *y, int32_t B, int32_t H, int32_t W,
                       int32_t f, float *x, int32_t accumulate, void *stream);
/* Full-weighting (tent) restriction by f and its adjoint (bilinear
 * prolongation): coarse point I is centred on fine position f I + (f-1)/2,
 * fine pixel i weighs max(0, 2f - |2i - (2fI + f - 1)|) / (2 f^2) per axis. */
int mk_restrict_tent_f32(const float *x, int32_t B, int32_t H, int32_t W,
                         int32_t f, float *y, void *stream);
int mk_prolong_tent_f32(const float *y, int32_t B, int32_t H, int32_t W,
                        int32_t f, float *x, int32_t accumulate, void *stream);
/* z = s * (exp(log_alpha[0]) * u + g); g may be NULL (Jacobi-like). */
int mk_precon_finish_f32(float *z, const float *s, const float *log_alpha,
                         const float *u, const float *g, int64_t n,
                         void *stream);
/* Training loss L = sum m (s (alpha u + g) - e)^2 with q = dL/d(alpha u + g)
 * written per pixel and sc[0] = L, sc[1] = dL/dalpha (FP64, device). */
int64_t mk_precon_loss_workspace_bytes(void);
int mk_precon_loss_f32(const float *g, const float *u, const float *s,
                       const float *e, const float *m, const float *log_alpha,
                       int64_t n, float *q, double *sc, void *work,
                       int64_t work_bytes, void *stream);
/* Energy-norm training loss: z = M r (full map), az = A z from the mapmaking
 * operator, e the error with A e = r.  sc[0] = (z - e).(az - r)
 * = ||z - e||_A^2, q = 2 s (az - r) (the gradient with respect to
 * t = alpha u + K^T K u, z = s t), sc[1] = sum q u = dL/dalpha. */
int mk_precon_aloss_f32(const float *z, const float *az, const float *r,
                        const float *e, const float *s, const float *u,
                        int64_t n, float *q, double *sc, void *work,
                        int64_t work_bytes, void *stream);
/* Adam step on a flat parameter vector whose last element is log(alpha);
 * that element's gradient is alpha * dalpha[0].  gscale multiplies every
 * gradient. */
int mk_adam_f32(float *p, const float *g, float *m, float *v, int32_t n,
                const double *dalpha, float lr, float beta1, float beta2,
                float eps, int32_t step, float gscale, void *stream);

/* ---- TOD preprocessing (examples/minkasi_mpi_example.py:42-52) ----------
 * fit_cm_plus_poly (minkasi/tods/processing.py:510-579) in device pieces.
 * Per column c (sample) of a row-major nrow x ncol block (row stride ld):
 * out[c] = np.median over the rows (the mean of the two middle values for an
 * even count); 1 <= nrow <= 1024.  Replaces np.median(dd, axis=0)
 * (processing.py:568). */
int mk_median_cols_f32(const float *dat, int64_t nrow, int64_t ncol,
                       int64_t ld, float *out, void *stream);
int mk_median_cols_f64(const double *dat, int64_t nrow, int64_t ncol,
                       int64_t ld, double *out, void *stream);
/* Design of the fit (processing.py:554-571): with x_t = (t - (n-1)/2) /
 * ((n-1)/2), X[t][k] = P_k(x_t) for k <= ord (legvander(xx, ord)) and
 * X[t][ord+1+i] = P_i(x_t) cm[t] for i < cm_ord (cm_pmat * cm); X is n x ncol
 * row-major and Xt its transpose, FP64, ncol = ord + 1 + cm_ord <= 8. */
int mk_cm_poly_design_f32(const float *cm, int64_t n, int ord, int cm_ord,
                          double *X, double *Xt, void *stream);
int mk_cm_poly_design_f64(const double *cm, int64_t n, int ord, int cm_ord,
                          double *X, double *Xt, void *stream);
/* D[k][j] = sum_t X[t][k] y[j][t] (ncol x nrow, FP64, fixed-order sums):
 * the normal-equation pieces mm^T mm and mm^T dat^T of _linfit_2mat
 * (processing.py:28-39).  work: mk_design_dots_workspace_bytes bytes. */
int64_t mk_design_dots_workspace_bytes(int64_t n, int ncol, int64_t nrow);
int mk_design_dots_f32(const double *X, int64_t n, int ncol, const float *y,
                       int64_t nrow, int64_t ld, double *D, void *work,
                       int64_t work_bytes, void *stream);
int mk_design_dots_f64(const double *X, int64_t n, int ncol, const double *y,
                       int64_t nrow, int64_t ld, double *D, void *work,
                       int64_t work_bytes, void *stream);
/* out[j][t] = base[j][t] - sum_{k<nk} X[t][c0+k] coef[k][j], or the sum
 * alone when base is NULL: dd = dat - pmat.fitp_p and pred2 = cm_mat.fitp_cm
 * (processing.py:573-575); coef is nk x nrow. */
int mk_design_apply_f32(const float *base, int64_t nrow, int64_t n, int64_t ld,
                        const double *X, int ncol, int c0, int nk,
                        const double *coef, float *out, int64_t ld_out,
                        void *stream);
int mk_design_apply_f64(const double *base, int64_t nrow, int64_t n,
                        int64_t ld, const double *X, int ncol, int c0, int nk,
                        const double *coef, double *out, int64_t ld_out,
                        void *stream);

/* ---- Cut samples (CutsCompact, minkasi/tods/cuts.py:201-498) -------------
 * imap: device int64[ncut], flat det * ndata + sample indices into a TOD of
 * nsamp samples.  Indices outside [0, nsamp) are skipped and counted into
 * *dev_nbad (device int64, may be NULL).
 * vec[i] (+)= dat[imap[i]].  Replaces tod2cuts (minkasi.c:260-268). */
int mk_tod2cuts_f32(float *vec, const float *dat, int64_t nsamp,
                    const int64_t *imap, int64_t ncut, int do_add,
                    int64_t *dev_nbad, void *stream);
int mk_tod2cuts_f64(double *vec, const double *dat, int64_t nsamp,
                    const int64_t *imap, int64_t ncut, int do_add,
                    int64_t *dev_nbad, void *stream);
/* dat[imap[i]] (+)= vec[i].  imap must not repeat an index (the reference's
 * serial loop would apply repeats in order; here they would race).
 * Replaces cuts2tod (minkasi.c:271-284). */
int mk_cuts2tod_f32(float *dat, const float *vec, int64_t nsamp,
                    const int64_t *imap, int64_t ncut, int do_add,
                    int64_t *dev_nbad, void *stream);
int mk_cuts2tod_f64(double *dat, const double *vec, int64_t nsamp,
                    const int64_t *imap, int64_t ncut, int do_add,
                    int64_t *dev_nbad, void *stream);

/* ---- Polarised maps (PolMap, minkasi/maps/polmap.py:47-724) --------------
 * A map holds ncomp values per pixel, index ncomp * ipix + c (the
 * reference's [nx, ny, npol] array).  twogamma: device array, the
 * polarisation angle 2 gamma of every sample, in the map's precision.
 * Pixels outside [0, npix) are skipped and counted into *dev_nbad. */
#define MK_POL_QU 1         /* (Q, U)                         ncomp 2 */
#define MK_POL_IQU 2        /* (I, Q, U)                      ncomp 3 */
#define MK_POL_QU_PRECON 3  /* (QQ, UU, QU)                   ncomp 3 */
#define MK_POL_IQU_PRECON 4 /* (I, Q, U, QQ, QU, UU)          ncomp 6 */
/* npol 3: dat (+)= m[3p] + m[3p+1] cos 2g + m[3p+2] sin 2g; npol 2:
 * dat (+)= m[2p] cos 2g + m[2p+1] sin 2g.  Replaces map2tod_iqu_omp and
 * map2tod_qu_omp (minkasi.c:146-175) behind polmap2tod
 * (mapmaking/map2tod.py:57-120). */
int mk_polmap2tod_f32(float *dat, const float *map, const float *twogamma,
                      int64_t nsamp, const int32_t *ipix, int64_t npix,
                      int npol, int do_add, int64_t *dev_nbad, void *stream);
int mk_polmap2tod_f64(double *dat, const double *map, const double *twogamma,
                      int64_t nsamp, const int32_t *ipix, int64_t npix,
                      int npol, int do_add, int64_t *dev_nbad, void *stream);
/* map += the per-sample terms of mode (MK_POL_*) with hardware atomics
 * (summation order not fixed).  Replaces tod2map_qu_simple,
 * tod2map_iqu_simple, tod2map_qu_precon_simple and
 * tod2map_iqu_precon_simple (minkasi.c:177-257) behind tod2polmap
 * (mapmaking/tod2map.py:164-215). */
int mk_tod2polmap_atomic_f32(float *map, const float *dat,
                             const float *twogamma, int64_t nsamp,
                             const int32_t *ipix, int64_t npix, int mode,
                             int64_t *dev_nbad, void *stream);
int mk_tod2polmap_atomic_f64(double *map, const double *dat,
                             const double *twogamma, int64_t nsamp,
                             const int32_t *ipix, int64_t npix, int mode,
                             int64_t *dev_nbad, void *stream);
/* In place, every pixel's QU_PRECON 2x2 or IQU_PRECON 3x3 block replaced by
 * its pseudo-inverse (np.linalg.pinv(block, thresh)), computed in FP64.
 * IQU_PRECON also zeroes blocks whose determinant is below thresh * (the
 * map's largest determinant) and sets 1/I where such a block has I > 0.
 * work: 8 bytes of device memory (IQU_PRECON only; may be NULL for QU).
 * Replaces PolMap.invert (polmap.py:300-384). */
int mk_polmap_invert_f32(float *map, int64_t npix, int mode, double thresh,
                         void *work, int64_t work_bytes, void *stream);
int mk_polmap_invert_f64(double *map, int64_t npix, int mode, double thresh,
                         void *work, int64_t work_bytes, void *stream);
/* out = block(pre) . x per pixel: pre is a QU_PRECON (x, out: QU) or
 * IQU_PRECON (x, out: IQU) map.  out may equal x.  Replaces
 * PolMap.__mul__ (polmap.py:624-684). */
int mk_polmap_precon_mul_f32(float *out, const float *pre, const float *x,
                             int64_t npix, int mode, void *stream);
int mk_polmap_precon_mul_f64(double *out, const double *pre, const double *x,
                             int64_t npix, int mode, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MKGPU_H */
