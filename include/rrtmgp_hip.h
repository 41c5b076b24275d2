/*
 * rrtmgp_hip.h -- C ABI of the MI355X (gfx950) RTE+RRTMGP column-radiation kernels.
 *
 * This is the thin layer the C++ class API (Gas_optics_rrtmgp_gpu, Cloud_optics_gpu,
 * Rte_lw_gpu, Rte_sw_gpu, Fluxes_*_gpu) calls, and the surface an FFI binds.  It replaces the
 * reference's two internal kernel boundaries:
 *   - the CUDA launcher namespaces of the include_kernels_cuda headers (device pointers), and
 *   - the Fortran C-ABI of include/rrtmgp_kernels.h:31-200 (CPU path).
 * Conventions (all entry points):
 *   - arrays are device pointers in the reference's Fortran layouts (first index fastest,
 *     e.g. tau(ncol, nlay, ngpt) has idx = icol + ncol*(ilay + nlay*igpt));
 *   - index tables (band_lims_gpt, gpoint_flavor, flavor, ...) hold 1-based values as in the
 *     reference; Bool is signed char (RTE_USE_CBOOL);
 *   - extents are int64_t; every array offset is computed in 64 bits;
 *   - stream is a hipStream_t (NULL = default stream); launches are asynchronous;
 *   - the return value is RRTMGP_OK or an error code; rrtmgp_hip_last_error() gives the
 *     message of the last failure on the calling thread;
 *   - _f32 / _f64 suffixes select single / double precision of the real arrays.
 */
#ifndef RRTMGP_HIP_H
#define RRTMGP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum
{
    RRTMGP_OK = 0,
    RRTMGP_ERR_ARG = 1,      /* invalid argument (std::invalid_argument / runtime_error in C++) */
    RRTMGP_ERR_RANGE = 2,    /* input out of the table range (std::range_error in C++)          */
    RRTMGP_ERR_HIP = 3,      /* HIP runtime error                                               */
    RRTMGP_ERR_RUNTIME = 4   /* any other failure                                               */
};

#define RRTMGP_MAX_GAS 32

typedef signed char rrtmgp_bool;

/* ------------------------------------------------------------------------------------------ */
/* Runtime helpers (FFI callers without their own HIP binding).  No reference counterpart:     */
/* the reference allocates through Tools_gpu::allocate_gpu (include/tools_gpu.h:72-99).        */
/* ------------------------------------------------------------------------------------------ */
const char* rrtmgp_hip_last_error(void);
int rrtmgp_hip_device_count(int* n);
int rrtmgp_hip_set_device(int device);
int rrtmgp_hip_get_device(int* device);
int rrtmgp_hip_device_name(char* buf, int len);
int rrtmgp_hip_malloc(void** ptr, size_t bytes);
int rrtmgp_hip_free(void* ptr);
int rrtmgp_hip_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int rrtmgp_hip_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int rrtmgp_hip_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int rrtmgp_hip_memset(void* ptr, int value, size_t bytes, void* stream);
int rrtmgp_hip_stream_create(void** stream);
int rrtmgp_hip_stream_destroy(void* stream);
int rrtmgp_hip_stream_synchronize(void* stream);
int rrtmgp_hip_device_synchronize(void);
int rrtmgp_hip_event_create(void** event);
int rrtmgp_hip_event_destroy(void* event);
int rrtmgp_hip_event_record(void* event, void* stream);
int rrtmgp_hip_event_elapsed_ms(void* start, void* stop, float* ms);

/* Caching device allocator (the reference's counterparts: allocate_gpu/free_gpu with its own */
/* pool, include/tools_gpu.h:71-98, Memory_pool_gpu include/mem_pool_gpu.h:12, src_cuda/mem_pool_gpu.cu). */
/* A block remembers the stream it was allocated on.  Freed on that stream, it is handed out  */
/* again only to a request on the same device and stream, so stream order covers reuse.  Freed */
/* on another stream, it is fenced by an event on each of the two streams and a later request  */
/* on any stream of the device waits on both before using it.  Streams given to the library    */
/* must outlive their blocks or be destroyed with rrtmgp_hip_stream_destroy.  release()        */
/* synchronizes and frees all unused blocks.                                                   */
int rrtmgp_hip_cache_malloc(void** ptr, size_t bytes, void* stream);
int rrtmgp_hip_cache_free(void* ptr, void* stream);
int rrtmgp_hip_cache_release(void);
int rrtmgp_hip_cache_stats(size_t* live_bytes, size_t* cached_bytes);

/* Array_gpu element operations (reference include/Array.h): fill sets n elements of          */
/* elem_size bytes (1, 2, 4 or 8) to *value (a host pointer), as Array_gpu::fill and fill_kernel*/
/* (:296-335, :501-511).  subset copies dst(d_1..d_n) = src(s_1..s_n), s_k = d_k + starts_k - 1 */
/* (1-based starts, relative to src_offsets when given) or s_k = 1 where src_dims_k == 1 (the  */
/* spread of Array_gpu::subset, :563-607, subset_kernel :296-322); ndim <= 8.  Ranges outside  */
/* the source are rejected with RRTMGP_ERR_ARG.                                                */
int rrtmgp_hip_fill(void* dst, int64_t n, int elem_size, const void* value, void* stream);
int rrtmgp_hip_subset(int elem_size, int ndim, const void* src, const int64_t* src_dims,
                      const int64_t* src_offsets, const int64_t* starts, void* dst,
                      const int64_t* dst_dims, void* stream);

/* Per-kernel timing.  No counterpart in the reference (its timing is external profiling):   */
/* with timing on, every kernel launch of this library is bracketed by two HIP events on its */
/* launch stream.  enable(1) clears earlier records; collect() waits for the recorded events  */
/* and totals them per kernel name; get(i) reads total i (kernel name, launches, total ms).   */
int rrtmgp_hip_timing_enable(int on);
int rrtmgp_hip_timing_collect(int* n_kernels);
int rrtmgp_hip_timing_get(int i, char* name, int len, int64_t* launches, double* total_ms);

/* ------------------------------------------------------------------------------------------ */
/* Device-resident k-distribution, as prepared by the Gas_optics_rrtmgp_gpu constructor        */
/* (reference: src_cuda/Gas_optics_rrtmgp.cu:779-802 uploads the same tables).  Real-valued    */
/* pointers hold float or double matching the entry-point suffix.                              */
/* ------------------------------------------------------------------------------------------ */
typedef struct
{
    int64_t ngas, nflav, neta, npres, ntemp, ngpt, nbnd, nplanck;
    int64_t idx_h2o;                       /* 1-based index of h2o in the gas list          */
    double press_ref_log_delta, temp_ref_min, temp_ref_delta, press_ref_trop_log;
    double totplnk_delta;
    int64_t nminor_lower, nminor_upper;
    const int* flavor;                     /* (2, nflav)                                    */
    const int* gpoint_flavor;              /* (2, ngpt)                                     */
    const int* band_lims_gpt;              /* (2, nbnd)                                     */
    const void* press_ref_log;             /* (npres)                                       */
    const void* temp_ref;                  /* (ntemp)                                       */
    const void* vmr_ref;                   /* (2, ngas+1, ntemp), gas 0 = dry air           */
    const void* kmajor;                    /* (ntemp, neta, npres+1, ngpt)                  */
    const void* kminor_lower;              /* (ntemp, neta, ncontrib_lower)                 */
    const void* kminor_upper;
    const int* minor_limits_gpt_lower;     /* (2, nminor)                                   */
    const int* minor_limits_gpt_upper;
    const rrtmgp_bool* minor_scales_with_density_lower;
    const rrtmgp_bool* minor_scales_with_density_upper;
    const rrtmgp_bool* scale_by_complement_lower;
    const rrtmgp_bool* scale_by_complement_upper;
    const int* idx_minor_lower;            /* 1-based gas index                             */
    const int* idx_minor_upper;
    const int* idx_minor_scaling_lower;    /* 1-based gas index, <= 0 for none              */
    const int* idx_minor_scaling_upper;
    const int* kminor_start_lower;         /* 1-based contributor offset                    */
    const int* kminor_start_upper;
    const void* totplnk;                   /* (nplanck, nbnd)          LW only              */
    const void* planck_frac;               /* (ntemp, neta, npres+1, ngpt)  LW only         */
    const void* krayl;                     /* (ntemp, neta, ngpt, 2)   SW only              */
    int64_t ncontrib_lower, ncontrib_upper;  /* last dims of kminor_lower / kminor_upper     */
    /* Band-packed tables of the fused gas-optics entry points, filled in by               */
    /* rrtmgp_hip_kdist_pack_* from the tables above (NULL until then).  Per band, the       */
    /* band's g-points (padded with zeros to a multiple of 4) are the fastest index.         */
    const void* go_bands;                  /* per-band descriptors (library-internal)       */
    const void* kmajor_p;                  /* per band (gpad, npres+1, neta, ntemp)         */
    const void* planck_frac_p;             /* as kmajor_p              LW only              */
    const void* kminor_p;                  /* per band (gpad, slot, neta, ntemp, half)      */
    const void* krayl_p;                   /* per band (gpad, neta, ntemp, half)  SW only   */
    int64_t max_band_gpts;                 /* largest number of g-points in one band        */
} rrtmgp_kdist_dev;

#define RRTMGP_DECLARE_BOTH(DECL) DECL(f32, float) DECL(f64, double)

/* Builds the band-packed tables of a k-distribution whose reference-layout tables (above) */
/* are filled in, and stores their device pointers in *kd.  No counterpart in the           */
/* reference: its kernels gather from the reference layout                                  */
/* (src_cuda/Gas_optics_rrtmgp.cu:779-802 uploads only that).  Requires every g-point of a  */
/* band to share the band's flavor in each half of the atmosphere, at most 8 minor          */
/* intervals per band and half, and each interval to have its band's flavor.  *pack owns the */
/* device memory until rrtmgp_hip_kdist_pack_free.  Synchronous (one-off set-up).            */
#define RRTMGP_DECL_KDIST_PACK(S, T)                                                           \
    int rrtmgp_hip_kdist_pack_##S(rrtmgp_kdist_dev* kd, void** pack);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_KDIST_PACK)
int rrtmgp_hip_kdist_pack_free(void* pack);

/* Volume mixing ratios in the k-distribution's gas order; dim1 is 1 or ncol, dim2 is 1 or   */
/* nlay (the broadcast rules of Gas_optics_rrtmgp::compute_gas_taus,                          */
/* src/Gas_optics_rrtmgp.cpp:1121-1154).                                                     */
typedef struct
{
    int64_t ngas;
    const void* vmr[RRTMGP_MAX_GAS];
    int64_t dim1[RRTMGP_MAX_GAS];
    int64_t dim2[RRTMGP_MAX_GAS];
} rrtmgp_gas_vmr;


/* ------------------------------------------------------------------------------------------ */
/* Gas optics                                                                                  */
/* ------------------------------------------------------------------------------------------ */

/* Fused gas optics, longwave: col_gas fill + interpolation + major/minor absorption + Planck  */
/* sources in one pass, optionally adding band-resolved cloud absorption (add_to of a 1scl).   */
/* Replaces Gas_optics_rrtmgp_gpu::gas_optics LW (src_cuda/Gas_optics_rrtmgp.cu:907-943):     */
/* fill_gases_kernel, interpolation, zero_array, compute_tau_absorption, compute_planck_source */
/* (+ inc_1scalar_by_1scalar_bybnd when cloud_tau != NULL).  sfc_lay is 1-based               */
/* (src/Gas_optics_rrtmgp.cpp:1333).                                                          */
#define RRTMGP_DECL_GAS_OPTICS_LW(S, T)                                                        \
    int rrtmgp_hip_gas_optics_lw_##S(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay,  \
        const T* play, const T* tlay, const T* tlev, const T* tsfc, const T* col_dry,          \
        const rrtmgp_gas_vmr* gases, int sfc_lay, const T* cloud_tau,                          \
        T* tau, T* lay_src, T* lev_src_inc, T* lev_src_dec, T* sfc_src, T* sfc_src_jac,        \
        void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_GAS_OPTICS_LW)

/* Fused gas optics, shortwave: absorption + Rayleigh + combine_abs_and_rayleigh (CPU        */
/* semantics: ssa = tau_ray/tau if tau > 2*epsilon, src/Gas_optics_rrtmgp.cpp:365-385), g = 0, */
/* optionally followed by inc_2stream_by_2stream_bybnd of band cloud properties.  Replaces     */
/* Gas_optics_rrtmgp_gpu::gas_optics SW (src_cuda/Gas_optics_rrtmgp.cu:947-977).             */
#define RRTMGP_DECL_GAS_OPTICS_SW(S, T)                                                        \
    int rrtmgp_hip_gas_optics_sw_##S(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay,  \
        const T* play, const T* tlay, const T* col_dry, const rrtmgp_gas_vmr* gases,           \
        const T* cloud_tau, const T* cloud_ssa, const T* cloud_g,                              \
        T* tau, T* ssa, T* g, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_GAS_OPTICS_SW)

/* Co-albedo forms (no reference counterpart).  In FP32, ssa near 1 resolves 1 - ssa only to     */
/* 6e-8, while the absorption of a Rayleigh-dominated layer at the model top is a few 1e-9 of its */
/* extinction: stored as ssa it is lost or multiplied, and the heating of the top layers with it. */
/* These entries carry the co-albedo 1 - ssa in place of ssa, formed from absorption optical     */
/* depths, through the shortwave chain: gas optics (+ band cloud add_to), cloud optics           */
/* (rrtmgp_hip_cloud_optics_* with two_stream = 2), aerosol optics (coalbedo = 1), the band      */
/* add_to (rrtmgp_hip_inc_bybnd_coalbedo_*) and the solver                                      */
/* (rrtmgp_hip_sw_solver_2stream_bnd_coalbedo_*).  Arguments are those of the ssa entries, with  */
/* co-albedos where those take ssa; results are the same quantities.                            */
#define RRTMGP_DECL_GAS_OPTICS_SW_CO(S, T)                                                     \
    int rrtmgp_hip_gas_optics_sw_coalbedo_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,        \
        int64_t nlay, const T* play, const T* tlay, const T* col_dry,                            \
        const rrtmgp_gas_vmr* gases, const T* cloud_tau, const T* cloud_coalbedo,               \
        const T* cloud_g, T* tau, T* coalbedo, T* g, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_GAS_OPTICS_SW_CO)

/* The fused gas optics restricted to listed (64-column tile, layer) pairs: tiles[i] =         */
/* ilay * ((ncol + 63)/64) + tile, 0-based, for i < ntiles (ntiles = 0 computes nothing);      */
/* outputs outside the listed pairs are left as they are.  No reference counterpart: the       */
/* fallback of the MLP surrogate (rrtmgp_hip_mlp_gas_optics_*) for pairs outside its training  */
/* envelope, whose tile list it writes.                                                        */
#define RRTMGP_DECL_GAS_OPTICS_TILES(S, T)                                                     \
    int rrtmgp_hip_gas_optics_lw_tiles_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,          \
        int64_t nlay, const int* tiles, int64_t ntiles, const T* play, const T* tlay,          \
        const T* tlev, const T* tsfc, const T* col_dry, const rrtmgp_gas_vmr* gases,           \
        int sfc_lay, const T* cloud_tau, T* tau, T* lay_src, T* lev_src_inc, T* lev_src_dec,   \
        T* sfc_src, T* sfc_src_jac, void* stream);                                             \
    int rrtmgp_hip_gas_optics_sw_tiles_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,          \
        int64_t nlay, const int* tiles, int64_t ntiles, const T* play, const T* tlay,          \
        const T* col_dry, const rrtmgp_gas_vmr* gases, const T* cloud_tau, const T* cloud_ssa, \
        const T* cloud_g, T* tau, T* ssa, T* g, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_GAS_OPTICS_TILES)

/* The fused entries above restricted to bands band_start .. band_start+band_count-1 (0-based):*/
/* the output g-point dimension is that range's g-points only (ngpt_range, counted from the    */
/* first g-point of band_start); cloud arrays stay (ncol, nlay, nbnd).  No counterpart in the  */
/* reference (whose gas optics always covers every band); used by the adjoint driver to        */
/* recompute gas optics band by band.                                                           */
#define RRTMGP_DECL_GAS_OPTICS_BANDS(S, T)                                                     \
    int rrtmgp_hip_gas_optics_lw_bands_##S(const rrtmgp_kdist_dev* kd, int64_t band_start,     \
        int64_t band_count, int64_t ncol, int64_t nlay, const T* play, const T* tlay,          \
        const T* tlev, const T* tsfc, const T* col_dry, const rrtmgp_gas_vmr* gases,           \
        int sfc_lay, const T* cloud_tau, T* tau, T* lay_src, T* lev_src_inc, T* lev_src_dec,   \
        T* sfc_src, T* sfc_src_jac, void* stream);                                              \
    int rrtmgp_hip_gas_optics_sw_bands_##S(const rrtmgp_kdist_dev* kd, int64_t band_start,     \
        int64_t band_count, int64_t ncol, int64_t nlay, const T* play, const T* tlay,          \
        const T* col_dry, const rrtmgp_gas_vmr* gases, const T* cloud_tau, const T* cloud_ssa,  \
        const T* cloud_g, T* tau, T* ssa, T* g, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_GAS_OPTICS_BANDS)

/* Gas_optics_rrtmgp_kernels_cuda::interpolation (include_kernels_cuda/                       */
/* gas_optics_rrtmgp_kernels_cuda.h:42-61; Fortran rrtmgp_interpolation, rrtmgp_kernels.h:60-79)*/
/* Materializes jtemp, jpress, tropo (ncol,nlay), jeta (2,ncol,nlay,nflav), col_mix (2,...),  */
/* fminor (2,2,...), fmajor (2,2,2,...) from col_gas (ncol,nlay,ngas+1; gas 0 = col_dry).     */
#define RRTMGP_DECL_INTERPOLATION(S, T)                                                        \
    int rrtmgp_hip_interpolation_##S(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay,  \
        const T* play, const T* tlay, const T* col_gas, int* jtemp, T* fmajor, T* fminor,      \
        T* col_mix, rrtmgp_bool* tropo, int* jeta, int* jpress, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_INTERPOLATION)

/* Gas_optics_rrtmgp_kernels_cuda::compute_tau_absorption (gas_optics_rrtmgp_kernels_cuda.h:  */
/* 81-115; Fortran rrtmgp_compute_tau_absorption, rrtmgp_kernels.h:81-108): tau += major +   */
/* minor (lower set where tropo, upper set elsewhere) from materialized interpolation.         */
#define RRTMGP_DECL_TAU_ABS(S, T)                                                              \
    int rrtmgp_hip_compute_tau_absorption_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,       \
        int64_t nlay, const rrtmgp_bool* tropo, const T* col_mix, const T* fmajor,             \
        const T* fminor, const T* play, const T* tlay, const T* col_gas, const int* jeta,      \
        const int* jtemp, const int* jpress, T* tau, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_TAU_ABS)

/* Gas_optics_rrtmgp_kernels_cuda::compute_tau_rayleigh (gas_optics_rrtmgp_kernels_cuda.h:   */
/* 68-79; Fortran rrtmgp_compute_tau_rayleigh, rrtmgp_kernels.h:129-138).                     */
#define RRTMGP_DECL_TAU_RAY(S, T)                                                              \
    int rrtmgp_hip_compute_tau_rayleigh_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,         \
        int64_t nlay, const T* col_dry, const T* col_gas, const T* fminor, const int* jeta,    \
        const rrtmgp_bool* tropo, const int* jtemp, T* tau_rayleigh, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_TAU_RAY)

/* Gas_optics_rrtmgp_kernels_cuda::combine_abs_and_rayleigh (gas_optics_rrtmgp_kernels_cuda.h:*/
/* 63-66), with the CPU threshold 2*epsilon (src/Gas_optics_rrtmgp.cpp:365-385).              */
#define RRTMGP_DECL_COMBINE(S, T)                                                              \
    int rrtmgp_hip_combine_abs_and_rayleigh_##S(int64_t ncol, int64_t nlay, int64_t ngpt,     \
        const T* tau_abs, const T* tau_rayleigh, T* tau, T* ssa, T* g, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_COMBINE)

/* Gas_optics_rrtmgp_kernels_cuda::compute_planck_source (gas_optics_rrtmgp_kernels_cuda.h:   */
/* 117-131; Fortran rrtmgp_compute_Planck_source, rrtmgp_kernels.h:119-127), delta_Tsurf = 1.  */
#define RRTMGP_DECL_PLANCK(S, T)                                                               \
    int rrtmgp_hip_compute_planck_source_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,        \
        int64_t nlay, const T* tlay, const T* tlev, const T* tsfc, int sfc_lay,                \
        const T* fmajor, const int* jeta, const rrtmgp_bool* tropo, const int* jtemp,          \
        const int* jpress, T* sfc_src, T* lay_src, T* lev_src_inc, T* lev_src_dec,             \
        T* sfc_src_jac, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_PLANCK)

/* Gas_optics_rrtmgp_gpu::get_col_dry (src_cuda/Gas_optics_rrtmgp.cu:806-903;                */
/* CPU src/Gas_optics_rrtmgp.cpp:764-792).                                                     */
#define RRTMGP_DECL_COL_DRY(S, T)                                                              \
    int rrtmgp_hip_get_col_dry_##S(int64_t ncol, int64_t nlay, const T* vmr_h2o,             \
        int64_t h2o_dim1, int64_t h2o_dim2, const T* plev, T* col_dry, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_COL_DRY)

/* fill_gases_kernel (src_cuda/Gas_optics_rrtmgp.cu:392-422): col_gas(ncol,nlay,ngas+1) with  */
/* gas 0 = col_dry and gas i = vmr_i * col_dry.                                               */
#define RRTMGP_DECL_FILL_GASES(S, T)                                                           \
    int rrtmgp_hip_fill_col_gas_##S(int64_t ncol, int64_t nlay, const rrtmgp_gas_vmr* gases,  \
        const T* col_dry, T* col_gas, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_FILL_GASES)

/* ------------------------------------------------------------------------------------------ */
/* Cloud optics (Cloud_optics_gpu::cloud_optics, src_cuda/Cloud_optics.cu:181-329; CPU       */
/* src/Cloud_optics.cpp:72-232).  LUTs are (nsize, nbnd) device arrays (ice: the roughness    */
/* category the host class selected).  two_stream = 0 writes the 1scl absorption tau only;    */
/* two_stream = 2 writes the co-albedo 1 - ssa in place of ssa (see the co-albedo forms);     */
/* delta_scale != 0 applies delta_scale_2str_k to the 2str result in the same pass.           */
/* ------------------------------------------------------------------------------------------ */
#define RRTMGP_DECL_CLOUD(S, T)                                                                \
    int rrtmgp_hip_cloud_optics_##S(int64_t ncol, int64_t nlay, int64_t nbnd,                 \
        int64_t liq_nsteps, T liq_step_size, T radliq_lwr, const T* lut_extliq,                \
        const T* lut_ssaliq, const T* lut_asyliq, int64_t ice_nsteps, T ice_step_size,         \
        T radice_lwr, const T* lut_extice, const T* lut_ssaice, const T* lut_asyice,           \
        const T* clwp, const T* ciwp, const T* reliq, const T* reice, int two_stream,          \
        int delta_scale, T* tau, T* ssa, T* g, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_CLOUD)

/* Aerosol optics (Aerosol_optics_gpu::aerosol_optics, src_cuda/Aerosol_optics.cu; CPU        */
/* src/Aerosol_optics.cpp:26-224).  aermr holds the 11 mass mixing ratios aermr01..aermr11    */
/* (ngas = 11; dim1 1|ncol, dim2 1|nlay), rh (ncol, nlay), plev (ncol, nlay+1).  Tables as    */
/* the class holds them: rh_upper (nhum), mext/ssa/g_phobic (nbnd, nphobic >= 11),            */
/* mext/ssa/g_philic (nbnd, nhum, nphilic >= 5).  Species, table columns and summation order  */
/* of the CPU path; layer mass |dp| / 9.81 (the GPU path's; equal to the CPU path's for levels */
/* ordered from the surface up); rh above every class bound takes the last class (the         */
/* reference reads past the table).  Outputs tau, ssa, g (ncol, nlay, nbnd), delta-scaled     */
/* (rte_delta_scale_2str_k) when delta_scale != 0; coalbedo != 0 writes 1 - ssa in place of   */
/* ssa (see the co-albedo forms).                                                              */
#define RRTMGP_DECL_AEROSOL(S, T)                                                              \
    int rrtmgp_hip_aerosol_optics_##S(int64_t ncol, int64_t nlay, int64_t nbnd, int64_t nhum, \
        int64_t nphobic, int64_t nphilic, const rrtmgp_gas_vmr* aermr, const T* rh,            \
        const T* plev, const T* rh_upper, const T* mext_phobic, const T* ssa_phobic,           \
        const T* g_phobic, const T* mext_philic, const T* ssa_philic, const T* g_philic,       \
        int delta_scale, int coalbedo, T* tau, T* ssa, T* g, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_AEROSOL)

/* ------------------------------------------------------------------------------------------ */
/* Optical-property arithmetic (Optical_props_kernels_cuda, include_kernels_cuda/             */
/* optical_props_kernels_cuda.h:31-55; Fortran rte_increment_*, rte_inc_*_bybnd,              */
/* rte_delta_scale_2str_k, rrtmgp_kernels.h:177-199).  eps = 3*tiny.  ssa/g NULL = 1scl.      */
/* ------------------------------------------------------------------------------------------ */
#define RRTMGP_DECL_OPTPROPS(S, T)                                                             \
    int rrtmgp_hip_increment_##S(int64_t ncol, int64_t nlay, int64_t ngpt, T* tau_inout,      \
        T* ssa_inout, T* g_inout, const T* tau_in, const T* ssa_in, const T* g_in,             \
        void* stream);                                                                         \
    int rrtmgp_hip_inc_bybnd_##S(int64_t ncol, int64_t nlay, int64_t ngpt, T* tau_inout,      \
        T* ssa_inout, T* g_inout, const T* tau_in, const T* ssa_in, const T* g_in,             \
        int64_t nbnd, const int* band_lims_gpt, void* stream);                                 \
    int rrtmgp_hip_delta_scale_2str_##S(int64_t ncol, int64_t nlay, int64_t ngpt, T* tau,     \
        T* ssa, T* g, void* stream);                                                            \
    int rrtmgp_hip_inc_bybnd_coalbedo_##S(int64_t ncol, int64_t nlay, int64_t ngpt,            \
        T* tau_inout, T* coalbedo_inout, T* g_inout, const T* tau_in, const T* coalbedo_in,    \
        const T* g_in, int64_t nbnd, const int* band_lims_gpt, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_OPTPROPS)

/* ------------------------------------------------------------------------------------------ */
/* Solvers (Rte_solver_kernels_cuda, include_kernels_cuda/rte_solver_kernels_cuda.h:31-65;    */
/* Fortran rte_lw_solver_noscat / rte_sw_solver_2stream / apply_BC_*, rrtmgp_kernels.h:     */
/* 140-175).                                                                                   */
/* ------------------------------------------------------------------------------------------ */

/* apply_BC: flux_dn(top) = 0 | inc_flux | inc_flux * factor(icol).  inc_flux (ncol, ngpt).  */
#define RRTMGP_DECL_APPLY_BC(S, T)                                                             \
    int rrtmgp_hip_apply_bc_##S(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1,       \
        const T* inc_flux, const T* factor, T* flux_dn, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_APPLY_BC)

/* expand_and_transpose (src_cuda/Rte_lw.cu:36-56): (nbnd, ncol) -> (ncol, ngpt).             */
#define RRTMGP_DECL_EXPAND(S, T)                                                               \
    int rrtmgp_hip_expand_and_transpose_##S(int64_t ncol, int64_t nbnd, int64_t ngpt,         \
        const int* band_lims_gpt, const T* arr_in, T* arr_out, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_EXPAND)

/* lw_solver_noscat with Gaussian quadrature (rte_solver_kernels_cuda.h:47-58; CPU           */
/* src_kernels/mo_rte_solver_kernels.F90:58-376).  secants (ncol, ngpt, nmus), weights (nmus) */
/* are device arrays; sfc_emis, sfc_src, inc_flux (NULL = 0) are (ncol, ngpt).  With          */
/* do_broadband the spectral fluxes are not returned (flux_up/flux_dn may be NULL) and         */
/* broadband_up/dn (ncol, nlev) hold the g-point sums.  do_jacobians adds the broadband       */
/* surface-temperature Jacobian of flux_up into flux_up_jac (ncol, nlev).                      */
#define RRTMGP_DECL_LW_SOLVER(S, T)                                                            \
    int rrtmgp_hip_lw_solver_noscat_##S(int64_t ncol, int64_t nlay, int64_t ngpt, int top_at_1,\
        int64_t nmus, const T* secants, const T* weights, const T* tau, const T* lay_source,   \
        const T* lev_source_inc, const T* lev_source_dec, const T* sfc_emis,                   \
        const T* sfc_src, const T* inc_flux, T* flux_up, T* flux_dn, int do_broadband,         \
        T* broadband_up, T* broadband_dn, int do_jacobians, const T* sfc_src_jac,              \
        T* flux_up_jac, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_LW_SOLVER)

/* lw_secants_array (rte_solver_kernels_cuda.h:62-64): secants (ncol, ngpt, n_gauss_quad)    */
/* filled with gauss_Ds(imu, n_gauss_quad) of the device table (max_gauss_pts, max_gauss_pts).*/
#define RRTMGP_DECL_LW_SECANTS(S, T)                                                           \
    int rrtmgp_hip_lw_secants_array_##S(int64_t ncol, int64_t ngpt, int64_t n_gauss_quad,    \
        int64_t max_gauss_pts, const T* gauss_Ds, T* secants, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_LW_SECANTS)

/* sw_solver_2stream (rte_solver_kernels_cuda.h:36-44; CPU F90:512-605, 996-1227).  mu0      */
/* (ncol); sfc_alb_dir/dif, inc_flux_dir, inc_flux_dif (ncol, ngpt).  flux_dn is total        */
/* (diffuse + direct).  do_broadband as for the LW solver.                                     */
#define RRTMGP_DECL_SW_SOLVER(S, T)                                                            \
    int rrtmgp_hip_sw_solver_2stream_##S(int64_t ncol, int64_t nlay, int64_t ngpt,            \
        int top_at_1, const T* tau, const T* ssa, const T* g, const T* mu0,                    \
        const T* sfc_alb_dir, const T* sfc_alb_dif, const T* inc_flux_dir, T* flux_up,         \
        T* flux_dn, T* flux_dir, int has_dif_bc, const T* inc_flux_dif, int do_broadband,      \
        T* broadband_up, T* broadband_dn, T* broadband_dir, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_SW_SOLVER)

/* Band-summed variants (no counterpart in the reference, whose drivers reduce materialised   */
/* g-point fluxes with Fluxes_byband::reduce, src/Fluxes.cpp:142-197): the solvers sum the    */
/* fluxes over the g-points of each band (band_lims_gpt (2, nbnd), 1-based, host memory) and  */
/* write band fluxes (ncol, nlev, nbnd) and/or broadband fluxes (ncol, nlev); either set may  */
/* be NULL.  The SW flux_dn outputs are total (diffuse + direct).  They take the surface       */
/* inputs as the class API holds them, so nothing is expanded per call: sfc_emis and the       */
/* albedos are (nbnd, ncol) (the reference's Rte_lw/Rte_sw arguments, expanded there by        */
/* expand_and_transpose, src_cuda/Rte_lw.cu:36-56), and secants / weights are the nmus Gauss   */
/* secants and weights of the quadrature (device arrays of nmus values; the reference fills    */
/* (ncol, ngpt, nmus) secants with the same value per angle, rte_solver_kernels.cu:16-32).     */
/* Heating rates (no reference counterpart): when heating_rate (ncol, nlay) is given, with plev  */
/* (ncol, nlev), each layer's net-flux convergence is formed per g-point from the layer's own    */
/* terms -- LW: -fac (S - (1 - T) I_in) per sweep; SW: (dn_dif(top) + up(bottom)) (1 - R - T) + */
/* dir(top) (1 - Tn) - S_up - S_dn, with 1 - R - T and 1 - Tn formed without cancellation -- and */
/* summed over g-points in FP64; hr = g0/cp * convergence / |dp| [K/day], g0 = 9.80665 m s-2,    */
/* cp = 1004.64 J kg-1 K-1.  This is the difference g0/cp (net_i - net_i+1) / (p_i+1 - p_i) of  */
/* rrtmgp_hip_heating_rate_*, without differencing rounded level fluxes.  NULL: no heating rate.*/
#define RRTMGP_DECL_SOLVERS_BND(S, T)                                                          \
    int rrtmgp_hip_lw_solver_noscat_bnd_##S(int64_t ncol, int64_t nlay, int64_t ngpt,          \
        int top_at_1, int64_t nmus, const T* secants, const T* weights, const T* tau,           \
        const T* lay_source, const T* lev_source_inc, const T* lev_source_dec,                 \
        const T* sfc_emis, const T* sfc_src, const T* inc_flux, int64_t nbnd,                  \
        const int* band_lims_gpt, T* bnd_flux_up, T* bnd_flux_dn, T* broadband_up,             \
        T* broadband_dn, const T* plev, T* heating_rate, void* stream);                         \
    int rrtmgp_hip_sw_solver_2stream_bnd_##S(int64_t ncol, int64_t nlay, int64_t ngpt,         \
        int top_at_1, const T* tau, const T* ssa, const T* g, const T* mu0,                    \
        const T* sfc_alb_dir, const T* sfc_alb_dif, const T* inc_flux_dir, int has_dif_bc,     \
        const T* inc_flux_dif, int64_t nbnd, const int* band_lims_gpt, T* bnd_flux_up,         \
        T* bnd_flux_dn, T* bnd_flux_dir, T* broadband_up, T* broadband_dn, T* broadband_dir,   \
        const T* plev, T* heating_rate, void* stream);                                          \
    int rrtmgp_hip_sw_solver_2stream_bnd_coalbedo_##S(int64_t ncol, int64_t nlay,             \
        int64_t ngpt, int top_at_1, const T* tau, const T* coalbedo, const T* g, const T* mu0,  \
        const T* sfc_alb_dir, const T* sfc_alb_dif, const T* inc_flux_dir, int has_dif_bc,     \
        const T* inc_flux_dif, int64_t nbnd, const int* band_lims_gpt, T* bnd_flux_up,         \
        T* bnd_flux_dn, T* bnd_flux_dir, T* broadband_up, T* broadband_dn, T* broadband_dir,   \
        const T* plev, T* heating_rate, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_SOLVERS_BND)

/* rrtmgp_hip_sw_solver_2stream_bnd_* without allocation.  No counterpart in the reference: it  */
/* is for callers that cannot allocate device memory (code under HIP graph capture, serving     */
/* loops, harnesses that forbid hipMalloc).  Same arguments, kernels and results as the          */
/* allocating entry; its scratch (SW checkpoints, group partial fluxes, FP64 heating-rate parts) */
/* comes from a fixed 64 MiB device array that the library holds on each device for the life   */
/* of the process, and it calls no hipMalloc, hipFree or allocating routine.  A call whose       */
/* scratch exceeds 64 MiB returns RRTMGP_ERR_ARG before any launch (at 128 layers and 224       */
/* g-points in FP32, about 86 KB per column plus 14 KB with heating rates: ~650 columns per      */
/* call; half as many in FP64).                                                                 */
/* Calls that overlap on one device share the array and race: serialise them (one stream, or    */
/* events).  Every scratch element is written before it is read, so nothing needs clearing.     */
#define RRTMGP_DECL_SW_NOALLOC(S, T)                                                           \
    int rrtmgp_hip_sw_solver_2stream_bnd_noalloc_##S(int64_t ncol, int64_t nlay, int64_t ngpt, \
        int top_at_1, const T* tau, const T* ssa, const T* g, const T* mu0,                    \
        const T* sfc_alb_dir, const T* sfc_alb_dif, const T* inc_flux_dir, int has_dif_bc,     \
        const T* inc_flux_dif, int64_t nbnd, const int* band_lims_gpt, T* bnd_flux_up,         \
        T* bnd_flux_dn, T* bnd_flux_dir, T* broadband_up, T* broadband_dn, T* broadband_dir,   \
        const T* plev, T* heating_rate, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_SW_NOALLOC)

/* Longwave with factored Planck sources.  No counterpart in the reference, whose gas optics */
/* writes lay_source, lev_source_inc and lev_source_dec per g-point (Source_func_lw,          */
/* include/Source_functions.h:73-99) for Rte_lw to read.  Every one of them is a g-point's     */
/* Planck fraction times its band's Planck function at the layer or level.  The factored pair */
/* keeps that product unevaluated:                                                              */
/*   gas_optics_lw_factored   writes tau, planck_frac (ncol, nlay, ngpt), band_lay_src         */
/*                            (tiles / ntiles as rrtmgp_hip_gas_optics_lw_tiles_*, or NULL / 0 */
/*                            for every pair)                                                   */
/*                            (ncol, nlay, nbnd) = B_band(T_lay), band_lev_src (ncol, nlev,    */
/*                            nbnd) = B_band(T_lev), and sfc_src / sfc_src_jac as the plain    */
/*                            entry;                                                            */
/*   lw_solver_noscat_bnd_factored  reads them and forms lay_source = planck_frac band_lay_src,*/
/*                            lev_source_inc (layer i) = planck_frac band_lev_src (level i+1)  */
/*                            and lev_source_dec = planck_frac band_lev_src (level i), the same*/
/*                            FP products, so fluxes equal the plain pair's bit for bit.       */
/* Two arrays of (ncol, nlay, ngpt) are neither written nor read.  Used by the longwave driver */
/* (Radiation_solver_longwave_gpu), also with a surrogate (rrtmgp_hip_mlp_gas_optics_lw_factored */
/* and this entry's tile list for its fallback).  The solver's secants, weights and sfc_emis are */
/* those of rrtmgp_hip_lw_solver_noscat_bnd_*: per angle, and (nbnd, ncol).                     */
#define RRTMGP_DECL_LW_FACTORED(S, T)                                                          \
    int rrtmgp_hip_gas_optics_lw_factored_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,       \
        int64_t nlay, const int* tiles, int64_t ntiles, const T* play, const T* tlay,           \
        const T* tlev, const T* tsfc,                                                            \
        const T* col_dry, const rrtmgp_gas_vmr* gases, int sfc_lay, const T* cloud_tau,        \
        T* tau, T* planck_frac, T* band_lay_src, T* band_lev_src, T* sfc_src, T* sfc_src_jac,  \
        void* stream);                                                                          \
    int rrtmgp_hip_lw_solver_noscat_bnd_factored_##S(int64_t ncol, int64_t nlay, int64_t ngpt, \
        int top_at_1, int64_t nmus, const T* secants, const T* weights, const T* tau,           \
        const T* planck_frac, const T* band_lay_src, const T* band_lev_src, const T* sfc_emis,  \
        const T* sfc_src, const T* inc_flux, int64_t nbnd, const int* band_lims_gpt,           \
        T* bnd_flux_up, T* bnd_flux_dn, T* broadband_up, T* broadband_dn, const T* plev,       \
        T* heating_rate, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_LW_FACTORED)

/* Longwave solvers with scattering.                                                          */
/*   lw_solver_noscat_rescaled: rte_lw_solver_noscat with do_rescaling = 1 and its ssa, g      */
/*       arguments (include/rrtmgp_kernels.h:148-158; src_kernels/mo_rte_solver_kernels.F90    */
/*       :168-200 and lw_transport_1rescl :731-822, Tang et al. 2018).  Other arguments as      */
/*       rrtmgp_hip_lw_solver_noscat_*: secants (ncol, ngpt, nmus), sfc_emis, sfc_src and       */
/*       inc_flux (NULL = 0) (ncol, ngpt), spectral fluxes (ncol, nlev, ngpt) or, with          */
/*       do_broadband, broadband sums (ncol, nlev); the Jacobian needs do_broadband and is      */
/*       transported without the rescaling adjustment, as the reference's.                      */
/*   lw_solver_2stream: the two-stream longwave solver lw_solver_2stream (F90:386-453), whose   */
/*       argument list it keeps (lay_source is accepted and unused, as there): level sources   */
/*       combined by lw_combine_sources (:896-920), Fu et al. (1997) coefficients with the      */
/*       diffusivity secant 1.66 (:832-887), Toon et al. (1989) layer sources (:928-978) and    */
/*       the adding method with surface albedo 1 - sfc_emis (:1117-1227).  sfc_emis, sfc_src,   */
/*       inc_flux (NULL = 0, a flux, not a radiance) (ncol, ngpt); fluxes (ncol, nlev, ngpt).   */
/* Neither solver exists on the reference's GPU path.                                          */
#define RRTMGP_DECL_LW_SCATTERING(S, T)                                                        \
    int rrtmgp_hip_lw_solver_noscat_rescaled_##S(int64_t ncol, int64_t nlay, int64_t ngpt,    \
        int top_at_1, int64_t nmus, const T* secants, const T* weights, const T* tau,           \
        const T* ssa, const T* g, const T* lay_source, const T* lev_source_inc,                 \
        const T* lev_source_dec, const T* sfc_emis, const T* sfc_src, const T* inc_flux,        \
        T* flux_up, T* flux_dn, int do_broadband, T* broadband_up, T* broadband_dn,             \
        int do_jacobians, const T* sfc_src_jac, T* flux_up_jac, void* stream);                  \
    int rrtmgp_hip_lw_solver_2stream_##S(int64_t ncol, int64_t nlay, int64_t ngpt,            \
        int top_at_1, const T* tau, const T* ssa, const T* g, const T* lay_source,              \
        const T* lev_source_inc, const T* lev_source_dec, const T* sfc_emis, const T* sfc_src,  \
        const T* inc_flux, T* flux_up, T* flux_dn, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_LW_SCATTERING)

/* ------------------------------------------------------------------------------------------ */
/* Adjoints.  No counterpart in the reference (its LW path has only the surface-temperature   */
/* Jacobian, src/Rte_lw.cpp:140).                                                              */
/* ------------------------------------------------------------------------------------------ */

/* Reverse mode of rrtmgp_hip_lw_solver_noscat_* with broadband output: for cotangents         */
/* bb_up_cot, bb_dn_cot (ncol, nlev) of the broadband fluxes, writes the gradients of          */
/* L = sum(bb_up_cot * broadband_up + bb_dn_cot * broadband_dn) with respect to tau, lay_source,*/
/* lev_source_inc, lev_source_dec (ncol, nlay, ngpt) and sfc_src, sfc_emis (ncol, ngpt), for   */
/* the same forward inputs (inc_flux may be NULL; it receives no gradient).                    */
#define RRTMGP_DECL_LW_ADJOINT(S, T)                                                           \
    int rrtmgp_hip_lw_solver_noscat_adjoint_##S(int64_t ncol, int64_t nlay, int64_t ngpt,      \
        int top_at_1, int64_t nmus, const T* secants, const T* weights, const T* tau,           \
        const T* lay_source, const T* lev_source_inc, const T* lev_source_dec,                 \
        const T* sfc_emis, const T* sfc_src, const T* inc_flux, const T* bb_up_cot,            \
        const T* bb_dn_cot, T* d_tau, T* d_lay_source, T* d_lev_source_inc,                    \
        T* d_lev_source_dec, T* d_sfc_src, T* d_sfc_emis, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_LW_ADJOINT)

/* Reverse mode of rrtmgp_hip_sw_solver_2stream_* with broadband output: for cotangents        */
/* bb_up_cot, bb_dn_cot, bb_dir_cot (ncol, nlev) of the broadband up, total down and direct    */
/* fluxes, writes the gradients of L = sum(cot * flux) with respect to tau, ssa, g             */
/* (ncol, nlay, ngpt) and sfc_alb_dir, sfc_alb_dif (ncol, ngpt).  The layer properties are     */
/* differentiated in FP64 in both precisions.  Uses FP64 scratch of 14 (nlay+1) ngpt ncol      */
/* values; callers bound it by calling per band (or g-point range).                            */
#define RRTMGP_DECL_SW_ADJOINT(S, T)                                                           \
    int rrtmgp_hip_sw_solver_2stream_adjoint_##S(int64_t ncol, int64_t nlay, int64_t ngpt,     \
        int top_at_1, const T* tau, const T* ssa, const T* g, const T* mu0,                    \
        const T* sfc_alb_dir, const T* sfc_alb_dif, const T* inc_flux_dir,                     \
        const T* inc_flux_dif, const T* bb_up_cot, const T* bb_dn_cot, const T* bb_dir_cot,    \
        T* d_tau, T* d_ssa, T* d_g, T* d_sfc_alb_dir, T* d_sfc_alb_dif, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_SW_ADJOINT)

/* Reverse mode of the fused gas optics (rrtmgp_hip_gas_optics_{lw,sw}_bands_* for one band):  */
/* given the cotangents of the band's outputs ((ncol, nlay, ngb) with ngb = the band's         */
/* g-points; g_sfc_src (ncol, ngb)), ADDS the gradients with respect to the column amounts     */
/* d_col_gas (ncol, nlay, ngas+2: slot 0 dry air as a column, slots 1..ngas the gases, slot    */
/* ngas+1 the derivative with respect to col_dry at fixed volume mixing ratios, formed from the */
/* degree-1 homogeneity of every column quantity in col_dry), T_lay and ln p_lay               */
/* (ncol, nlay), and in the longwave T_lev at each layer's level ilay (d_tlev_lo) and          */
/* ilay+1 (d_tlev_hi) (ncol, nlay) and T_sfc (ncol).  The SW cotangents are those of the       */
/* outputs after the optional band cloud add_to.  Call band by band on one stream.             */
/* get_col_dry_adjoint: reverse mode of rrtmgp_hip_get_col_dry_* (vmr_h2o with its broadcast  */
/* dims); ADDS into d_vmr_h2o (ncol, nlay, full) and d_plev (ncol, nlev).                       */
#define RRTMGP_DECL_GO_ADJOINT(S, T)                                                           \
    int rrtmgp_hip_gas_optics_lw_adjoint_##S(const rrtmgp_kdist_dev* kd, int64_t band,         \
        int64_t ncol, int64_t nlay, const T* play, const T* tlay, const T* tlev, const T* tsfc, \
        const T* col_dry, const rrtmgp_gas_vmr* gases, int sfc_lay, const T* g_tau,           \
        const T* g_lay_src, const T* g_lev_src_inc, const T* g_lev_src_dec,                    \
        const T* g_sfc_src, T* d_col_gas, T* d_tlay, T* d_lnp_lay, T* d_tlev_lo,               \
        T* d_tlev_hi, T* d_tsfc, void* stream);                                                 \
    int rrtmgp_hip_gas_optics_sw_adjoint_##S(const rrtmgp_kdist_dev* kd, int64_t band,         \
        int64_t ncol, int64_t nlay, const T* play, const T* tlay, const T* col_dry,            \
        const rrtmgp_gas_vmr* gases, const T* cloud_tau, const T* cloud_ssa,                   \
        const T* cloud_g, const T* g_tau, const T* g_ssa, const T* g_g, T* d_col_gas,          \
        T* d_tlay, T* d_lnp_lay, void* stream);                                                 \
    int rrtmgp_hip_get_col_dry_adjoint_##S(int64_t ncol, int64_t nlay, const T* vmr_h2o,      \
        int64_t h2o_dim1, int64_t h2o_dim2, const T* plev, const T* g_col_dry, T* d_vmr_h2o,   \
        T* d_plev, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_GO_ADJOINT)

/* Assembly steps of the adjoint driver.  col_chain: from d_col_gas (ncol, nlay, ngas+2) to    */
/* d_vmr (ncol, nlay, ngas, overwritten; gases in the descriptor's order) and d_col_dry        */
/* (ncol, nlay, overwritten; slot ngas+1).  levels: d_tlev (ncol, nlev) from the per-layer     */
/* contributions                                                                               */
/* of rrtmgp_hip_gas_optics_lw_adjoint_*.  gpt_to_band: band sums of (ncol, ngpt) values into  */
/* (nbnd, ncol) (band_lims_gpt device, 1-based).                                               */
#define RRTMGP_DECL_ADJ_HELPERS(S, T)                                                          \
    int rrtmgp_hip_adjoint_col_chain_##S(int64_t ncol, int64_t nlay,                           \
        const rrtmgp_gas_vmr* gases, const T* col_dry, const T* d_col_gas, T* d_vmr,           \
        T* d_col_dry, void* stream);                                                            \
    int rrtmgp_hip_adjoint_levels_##S(int64_t ncol, int64_t nlay, const T* d_lo, const T* d_hi,\
        T* d_tlev, void* stream);                                                               \
    int rrtmgp_hip_gpt_to_band_##S(int64_t ncol, int64_t nbnd, const int* band_lims_gpt,      \
        const T* in, T* out, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_ADJ_HELPERS)

/* ------------------------------------------------------------------------------------------ */
/* Flux reductions (Fluxes_kernels_cuda, include_kernels_cuda/fluxes_kernels_cuda.h:31-50;    */
/* Fortran rte_sum_broadband / rte_net_broadband_precalc / sum_byband / net_byband_precalc,   */
/* rrtmgp_kernels.h:34-52).  By-band sums use the CPU semantics (1-based inclusive band_lims, */
/* mo_fluxes_byband_kernels.F90:22-79).                                                       */
/* ------------------------------------------------------------------------------------------ */
#define RRTMGP_DECL_FLUXES(S, T)                                                               \
    int rrtmgp_hip_sum_broadband_##S(int64_t ncol, int64_t nlev, int64_t ngpt,                \
        const T* gpt_flux, T* flux, void* stream);                                             \
    int rrtmgp_hip_net_broadband_##S(int64_t ncol, int64_t nlev, const T* flux_dn,            \
        const T* flux_up, T* flux_net, void* stream);                                          \
    int rrtmgp_hip_sum_byband_##S(int64_t ncol, int64_t nlev, int64_t ngpt, int64_t nbnd,     \
        const int* band_lims, const T* gpt_flux, T* bnd_flux, void* stream);                   \
    int rrtmgp_hip_net_byband_full_##S(int64_t ncol, int64_t nlev, int64_t ngpt,              \
        int64_t nbnd, const int* band_lims, const T* gpt_flux_dn, const T* gpt_flux_up,        \
        T* bnd_flux_net, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_FLUXES)

/* Heating rate, no reference counterpart: hr(ncol, nlay) [K/day] =                          */
/* g0/cp * (net(k) - net(k+1)) / (p(k+1) - p(k)) from flux_net = dn - up (ncol, nlev); the    */
/* flux and pressure differences are taken in double precision for both suffixes.             */
#define RRTMGP_DECL_HEATING(S, T)                                                              \
    int rrtmgp_hip_heating_rate_##S(int64_t ncol, int64_t nlay, const T* flux_up,             \
        const T* flux_dn, const T* plev, T* heating_rate, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_HEATING)

/* Reverse mode of rrtmgp_hip_heating_rate_*: for the cotangent g_heating_rate (ncol, nlay),   */
/* ADDS the gradients into g_flux_up, g_flux_dn and d_plev (ncol, nlev).  No reference         */
/* counterpart.                                                                                 */
#define RRTMGP_DECL_HEATING_ADJ(S, T)                                                          \
    int rrtmgp_hip_heating_rate_adjoint_##S(int64_t ncol, int64_t nlay, const T* flux_up,     \
        const T* flux_dn, const T* plev, const T* g_heating_rate, T* g_flux_up, T* g_flux_dn,  \
        T* d_plev, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_HEATING_ADJ)

/* ------------------------------------------------------------------------------------------ */
/* MLP gas-optics surrogate (FP32 path; BF16 matrix cores with FP32 accumulation).  No         */
/* counterpart in the reference.  Per (column, layer) the network maps x = [T, ln p, ln h2o,     */
/* ln o3, ln co2, ln ch4, ln n2o, f] (f = 1 where p > p_trop, the lower-atmosphere tables;      */
/* normalised with x_mean/x_std) through                                                        */
/* two softplus layers of 128 to n_out outputs (denormalised with y_mean/y_std), interleaved  */
/* per g-point: LW (ln(tau/col_dry), Planck fraction), SW (ln(tau_abs/col_dry),               */
/* ln(tau_ray/col_dry)).  Weights are row-major (out, in) BF16; W1 is (128, 16) with the 8    */
/* inputs first.  A (64-column tile, layer) unit with any input outside [x_lo, x_hi] goes on   */
/* the fallback list (1 + ntile nlay ints, ntile = (ncol + 63)/64): fallback[0] = number of     */
/* listed units, fallback[1 ..] = ilay * ntile + tile, the tile list of                         */
/* rrtmgp_hip_gas_optics_{lw,sw}_tiles_f32, which recomputes them with the plain gas optics.    */
/* ------------------------------------------------------------------------------------------ */
typedef struct
{
    const void* W1;                        /* (128, 16) BF16                                */
    const void* W2;                        /* (128, 128) BF16                               */
    const void* W3;                        /* (n_out, 128) BF16                             */
    const float* b1;
    const float* b2;
    const float* b3;
    const float* x_mean;                   /* (8)                                           */
    const float* x_std;
    const float* x_lo;                     /* training envelope of the raw inputs           */
    const float* x_hi;
    const float* y_mean;                   /* (n_out)                                       */
    const float* y_std;
    int64_t n_out;                         /* 2 ngpt, a multiple of 32                      */
    double p_trop;                         /* tropopause pressure of the k-distribution [Pa] */
} rrtmgp_mlp_net;

/* Inference, fused with the gas-optics epilogue (outputs as rrtmgp_hip_gas_optics_lw/sw_f32, */
/* including the band cloud add_to).  gases: exactly h2o, o3, co2, ch4, n2o in this order;    */
/* gpt2band (ngpt) 0-based band of each g-point; totplnk (nplanck, nbnd) FP32.                 */
int rrtmgp_hip_mlp_gas_optics_lw(const rrtmgp_mlp_net* net, int64_t ncol, int64_t nlay, int64_t ngpt,
                                 int64_t nbnd, const int* gpt2band, const float* play, const float* tlay,
                                 const float* tlev, const float* tsfc, const float* col_dry,
                                 const rrtmgp_gas_vmr* gases, const float* totplnk, int64_t nplanck,
                                 double temp_min, double dplanck, int sfc_lay, const float* cloud_tau,
                                 float* tau, float* lay_src, float* lev_src_inc, float* lev_src_dec,
                                 float* sfc_src, float* sfc_src_jac, int* fallback, void* stream);
/* The same with factored Planck sources (outputs as rrtmgp_hip_gas_optics_lw_factored_f32).  */
int rrtmgp_hip_mlp_gas_optics_lw_factored(const rrtmgp_mlp_net* net, int64_t ncol, int64_t nlay,
                                          int64_t ngpt, int64_t nbnd, const int* gpt2band, const float* play,
                                          const float* tlay, const float* tlev, const float* tsfc,
                                          const float* col_dry, const rrtmgp_gas_vmr* gases,
                                          const float* totplnk, int64_t nplanck, double temp_min,
                                          double dplanck, int sfc_lay, const float* cloud_tau, float* tau,
                                          float* planck_frac, float* band_lay_src, float* band_lev_src,
                                          float* sfc_src, float* sfc_src_jac, int* fallback, void* stream);
int rrtmgp_hip_mlp_gas_optics_sw(const rrtmgp_mlp_net* net, int64_t ncol, int64_t nlay, int64_t ngpt,
                                 int64_t nbnd, const int* gpt2band, const float* play, const float* tlay,
                                 const float* col_dry, const rrtmgp_gas_vmr* gases, const float* cloud_tau,
                                 const float* cloud_ssa, const float* cloud_g, float* tau, float* ssa,
                                 float* g, int* fallback, void* stream);

/* Training.  gemm_bf16: C[m sCm + n sCn] (+)= sum_k A[m sAm + k sAk] B[k sBk + n sBn], BF16      */
/* operands, FP32 result (accumulate != 0 adds to C).  Row-major (rows, n) arrays below.        */
int rrtmgp_hip_mlp_gemm_bf16(int64_t M, int64_t N, int64_t K, const void* A, int64_t sAm, int64_t sAk,
                             const void* B, int64_t sBk, int64_t sBn, float* C, int64_t sCm, int64_t sCn,
                             int accumulate, void* stream);
/* Z += bias (per column); H = softplus(Z) in BF16.                                             */
int rrtmgp_hip_mlp_bias_softplus(int64_t rows, int n, float* Z, const float* bias, void* H, void* stream);
/* Y += bias; dY = 2 (Y - T)/(rows n) in BF16 and FP32; loss_parts[block] = partial sums of      */
/* (Y - T)^2 over blocks of 256 elements.                                                        */
int rrtmgp_hip_mlp_mse_grad(int64_t rows, int n, float* Y, const float* bias, const float* T, void* dY,
                            float* dYf, float* loss_parts, void* stream);
/* dZ = dH * sigmoid(Z) (softplus'), BF16 and FP32.                                              */
int rrtmgp_hip_mlp_softplus_back(int64_t n, const float* dH, const float* Z, void* dZ, float* dZf,
                                 void* stream);
/* db[j] = sum over rows of dZ[r, j].                                                            */
int rrtmgp_hip_mlp_bias_grad(int64_t rows, int n, const float* dZ, float* db, void* stream);
/* Adam step `step` (1-based) on FP32 master weights p with moments m, v; writes the BF16 copy   */
/* p_bf16 when it is not NULL.                                                                   */
int rrtmgp_hip_mlp_adam(int64_t n, float* p, const float* g, float* m, float* v, void* p_bf16, float lr,
                        float beta1, float beta2, float eps, int64_t step, void* stream);
int rrtmgp_hip_mlp_to_bf16(int64_t n, const float* x, void* y, void* stream);
/* Training data from the plain gas optics: raw inputs X (ncol nlay, 16) with row = icol +      */
/* ncol ilay; targets Y (ncol nlay, 2 ngpt) from LW (a = tau, b = lay_src) or SW (a = tau,       */
/* b = ssa) outputs of a clear-sky run.                                                          */
int rrtmgp_hip_mlp_inputs(int64_t ncol, int64_t nlay, const float* play, const float* tlay,
                          const rrtmgp_gas_vmr* gases, double p_trop, float* X, void* stream);
int rrtmgp_hip_mlp_targets(int lw, int64_t ncol, int64_t nlay, int64_t ngpt, const int* gpt2band,
                           const float* tlay, const float* col_dry, const float* a, const float* b,
                           const float* totplnk, int64_t nplanck, double temp_min, double dplanck,
                           float* Y, void* stream);
/* dst[i, :] = src[idx[i], :] (rows of n); y = (x - mean)/std over the first n_norm columns     */
/* (others 0), BF16 and/or FP32.                                                                 */
int rrtmgp_hip_mlp_gather_rows(int64_t rows, int n, const int64_t* idx, const float* src, float* dst,
                               void* stream);
int rrtmgp_hip_mlp_normalize_rows(int64_t rows, int n, int n_norm, const float* x, const float* mean,
                                  const float* std_, void* y_bf16, float* y_f32, void* stream);

/* spread_col_kernel (src_cuda/Gas_optics_rrtmgp.cu:40-68): dst(icol, igpt) = src(igpt), and  */
/* the per-column scaling of the driver (toa_src * tsi_scaling, src_test/Radiation_solver.cpp: */
/* 739-743): arr(icol, j) *= factor(icol) for j < n2.                                          */
#define RRTMGP_DECL_SPREAD(S, T)                                                               \
    int rrtmgp_hip_spread_col_##S(int64_t ncol, int64_t ngpt, const T* src, T* dst,           \
        void* stream);                                                                         \
    int rrtmgp_hip_scale_by_col_##S(int64_t ncol, int64_t n2, T* arr, const T* factor,        \
        void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_SPREAD)

/* Column-block copy (the subset/get_from_subset kernels of the reference driver,             */
/* src_kernels_cuda/subset_kernels.cu:1-98, include/Array.h:296-335): for nrows rows of       */
/* (ncol, rows) arrays, dst(dst_col0 + i, r) = src(src_col0 + i, r) for i < ncopy.            */
#define RRTMGP_DECL_COPY_COLS(S, T)                                                            \
    int rrtmgp_hip_copy_cols_##S(const T* src, int64_t src_ncol, int64_t src_col0, T* dst,    \
        int64_t dst_ncol, int64_t dst_col0, int64_t ncopy, int64_t nrows, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_COPY_COLS)

/* Input validation of Gas_optics_rrtmgp::gas_optics (src/Gas_optics_rrtmgp.cpp:813-826):     */
/* count of values outside [lo, hi] (or below lo when hi is NaN), written to *count (host).    */
/* Synchronizes the stream.                                                                    */
#define RRTMGP_DECL_COUNT(S, T)                                                                \
    int rrtmgp_hip_count_outside_##S(int64_t n, const T* arr, T lo, T hi, int64_t* count,      \
        void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_COUNT)

/* All input checks of one Gas_optics_rrtmgp::gas_optics call (src/Gas_optics_rrtmgp.cpp:       */
/* 805-826) in one launch and one device-to-host copy: for each of n <= 8 device arrays of sizes  */
/* elements (host arrays of pointers, sizes and bounds), counts[k] = values outside [lo[k],      */
/* hi[k]] (below lo[k] when hi[k] is NaN; NaN counts as outside).  When play is given, also     */
/* play_ends = {play(1,1), play(1,nlay)} of a (ncol, nlay) array, from which the surface layer  */
/* is chosen (:1333).  Synchronizes the stream once (the reference's CPU path checks on the     */
/* host; its GPU path does not check, src_cuda/Gas_optics_rrtmgp.cu:907-943).                  */
#define RRTMGP_DECL_CHECK(S, T)                                                                \
    int rrtmgp_hip_check_ranges_##S(int n, const T* const* arrays, const int64_t* sizes,       \
        const T* lo, const T* hi, int64_t* counts, const T* play, int64_t ncol, int64_t nlay,  \
        T* play_ends, void* stream);
RRTMGP_DECLARE_BOTH(RRTMGP_DECL_CHECK)

#ifdef __cplusplus
}
#endif
#endif
