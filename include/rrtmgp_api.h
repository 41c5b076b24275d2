/*
 * rrtmgp_api.h -- C facade over the C++ class API (Gas_optics_rrtmgp_gpu, Cloud_optics_gpu,
 * Rte_lw_gpu, Rte_sw_gpu, Fluxes_*_gpu, Radiation_solver_*_gpu) for FFI callers.
 *
 * This is the binding a maintainer of the reference would add to reach its class API from
 * Python (the role of the stale Cython module python/radiation.pyx:63-222 of the reference):
 * opaque handles for the objects, device pointers for the arrays, an int status return and
 * rrtmgp_api_last_error() for the message of the exception the C++ layer threw
 * (RRTMGP_ERR_RANGE for std::range_error, RRTMGP_ERR_ARG for std::invalid_argument,
 * RRTMGP_ERR_RUNTIME otherwise).  librrtmgp_cpp_f64.so is built with Float = double,
 * librrtmgp_cpp_f32.so with Float = float (RTE_USE_SP); rrtmgp_real follows that choice.
 * Array layouts are the reference's (Fortran order, first index fastest).
 */
#ifndef RRTMGP_API_H
#define RRTMGP_API_H

#include <stdint.h>

#include "rrtmgp_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifdef RTE_USE_SP
typedef float rrtmgp_real;
#else
typedef double rrtmgp_real;
#endif

/* Coefficient tables as the reference reader hands them to the Gas_optics_rrtmgp constructors */
/* (src_test/Radiation_solver.cpp:70-284; include/Gas_optics_rrtmgp.h:50-121): host arrays in */
/* double precision, Fortran layouts, string lists '\n'-joined.                               */
typedef struct
{
    int ngas; const char* gas_names;
    int nbnd, ngpt, ntemp, npres, neta, nplanck;   /* nplanck = 0 for a shortwave table */
    const int* key_species;          /* (2, 2, nbnd)                */
    const int* band2gpt;             /* (2, nbnd)                   */
    const double* band_lims_wavenum; /* (2, nbnd)                   */
    const double* press_ref;         /* (npres)                     */
    double press_ref_trop;
    const double* temp_ref;          /* (ntemp)                     */
    double temp_ref_p, temp_ref_t;
    const double* vmr_ref;           /* (2, ngas+1, ntemp)          */
    const double* kmajor;            /* (ngpt, neta, npres+1, ntemp)*/
    int nminor_ident; const char* gas_minor; const char* identifier_minor;
    int nminor_lower, ncontrib_lower;
    const double* kminor_lower;      /* (ncontrib, neta, ntemp)     */
    const char* minor_gases_lower;
    const int* minor_limits_gpt_lower;               /* (2, nminor) */
    const signed char* minor_scales_with_density_lower;
    const char* scaling_gas_lower;
    const signed char* scale_by_complement_lower;
    const int* kminor_start_lower;
    int nminor_upper, ncontrib_upper;
    const double* kminor_upper;
    const char* minor_gases_upper;
    const int* minor_limits_gpt_upper;
    const signed char* minor_scales_with_density_upper;
    const char* scaling_gas_upper;
    const signed char* scale_by_complement_upper;
    const int* kminor_start_upper;
    const double* totplnk;           /* (nplanck, nbnd)             LW */
    const double* planck_frac;       /* (ngpt, neta, npres+1, ntemp) LW */
    const double* solar_src_quiet;   /* (ngpt)                      SW */
    const double* solar_src_facular;
    const double* solar_src_sunspot;
    double tsi_default, mg_default, sb_default;
    const double* rayl_lower;        /* (ngpt, neta, ntemp)         SW */
    const double* rayl_upper;
    int n_available; const char* available_gases;
} rrtmgp_kdist_file;

/* Cloud LUT as the Cloud_optics constructor takes it (src/Cloud_optics.cpp:29-69).           */
typedef struct
{
    int nbnd; const double* band_lims_wvn;      /* (2, nbnd) */
    double radliq_lwr, radliq_upr, radliq_fac, radice_lwr, radice_upr, radice_fac;
    int nsize_liq; const double* lut_extliq; const double* lut_ssaliq; const double* lut_asyliq;
    int nsize_ice; const double* lut_extice; const double* lut_ssaice; const double* lut_asyice;
} rrtmgp_cloud_lut;

/* Aerosol LUT as load_and_init_aerosol_optics reads it (src_test/Radiation_solver.cpp:329-366) */
/* for the Aerosol_optics constructor (include/Aerosol_optics.h:20-24).                        */
typedef struct
{
    int nbnd, nhum, nphobic, nphilic;
    const double* band_lims_wvn;     /* (2, nbnd)                                              */
    const double* rh_upper;          /* (nhum): relative_humidity2                             */
    const double* mext_phobic; const double* ssa_phobic; const double* g_phobic;   /* (nbnd, nphobic) */
    const double* mext_philic; const double* ssa_philic; const double* g_philic;   /* (nbnd, nhum, nphilic) */
} rrtmgp_aerosol_lut;

/* Gas concentrations: n named device arrays with dims (1|ncol, 1|nlay).                      */
typedef struct
{
    int n; const char* names; const void* const* data; const int64_t* dims;   /* dims: 2*n */
} rrtmgp_api_gases;

int rrtmgp_api_precision(void);
const char* rrtmgp_api_last_error(void);
int rrtmgp_api_set_stream(void* stream);

/* Gas_optics_rrtmgp_gpu (include/Gas_optics_rrtmgp.h:291-551). */
int rrtmgp_api_kdist_create(const rrtmgp_kdist_file* file, void** kdist);
int rrtmgp_api_kdist_destroy(void* kdist);
/* info: ngas nflav neta npres ntemp ngpt nbnd nplanck nminor_lower nminor_upper idx_h2o is_lw */
int rrtmgp_api_kdist_info(void* kdist, int64_t* info);
int rrtmgp_api_kdist_get_int(void* kdist, const char* name, int* out, int64_t cap, int64_t* n);
int rrtmgp_api_kdist_get_real(void* kdist, const char* name, double* out, int64_t cap, int64_t* n);
int rrtmgp_api_kdist_gas_names(void* kdist, char* buf, int len);
int rrtmgp_api_kdist_set_check_ranges(void* kdist, int on);
/* MLP gas-optics surrogate (no reference counterpart; FP32 library only): with a network set,  */
/* every gas optics of this k-distribution, including the solvers', evaluates it and          */
/* recomputes the units outside its training envelope with the plain kernels.  The network's   */
/* device arrays stay the caller's; net = NULL restores the plain gas optics.                 */
/* last_fallback: (64-column tile, layer) units the last surrogate call recomputed.            */
int rrtmgp_api_kdist_set_surrogate(void* kdist, const rrtmgp_mlp_net* net);
int rrtmgp_api_kdist_last_fallback(void* kdist, int64_t* n_units);

/* Gas_concs_gpu (include/Gas_concs.h:74-91). */
int rrtmgp_api_gas_concs_create(const rrtmgp_api_gases* gases, void** gas_concs);
int rrtmgp_api_gas_concs_destroy(void* gas_concs);

int rrtmgp_api_get_col_dry(int ncol, int nlay, const rrtmgp_real* vmr_h2o, int64_t h2o_dim1,
                           int64_t h2o_dim2, const rrtmgp_real* plev, rrtmgp_real* col_dry);

/* gas_optics LW / SW (include/Gas_optics_rrtmgp.h:389-408). */
int rrtmgp_api_gas_optics_lw(void* kdist, int ncol, int nlay, const rrtmgp_real* play,
                             const rrtmgp_real* plev, const rrtmgp_real* tlay,
                             const rrtmgp_real* tsfc, const rrtmgp_real* tlev, void* gas_concs,
                             const rrtmgp_real* col_dry, rrtmgp_real* tau, rrtmgp_real* lay_src,
                             rrtmgp_real* lev_src_inc, rrtmgp_real* lev_src_dec,
                             rrtmgp_real* sfc_src, rrtmgp_real* sfc_src_jac);
int rrtmgp_api_gas_optics_sw(void* kdist, int ncol, int nlay, const rrtmgp_real* play,
                             const rrtmgp_real* plev, const rrtmgp_real* tlay, void* gas_concs,
                             const rrtmgp_real* col_dry, rrtmgp_real* tau, rrtmgp_real* ssa,
                             rrtmgp_real* g, rrtmgp_real* toa_src);
/* As rrtmgp_api_gas_optics_sw, writing the co-albedo 1 - ssa in place of ssa (no reference      */
/* counterpart: the shortwave chain in co-albedo form, which FP32 needs to resolve absorption    */
/* below 6e-8 of extinction; not with a surrogate).                                              */
int rrtmgp_api_gas_optics_sw_coalbedo(void* kdist, int ncol, int nlay, const rrtmgp_real* play,
                                      const rrtmgp_real* plev, const rrtmgp_real* tlay, void* gas_concs,
                                      const rrtmgp_real* col_dry, rrtmgp_real* tau, rrtmgp_real* coalbedo,
                                      rrtmgp_real* g, rrtmgp_real* toa_src);

/* The materializing kernels of the reference launch chain on this k-distribution's device    */
/* tables: fill_gases_kernel (col_gas (ncol, nlay, ngas+1)) and interpolation (outputs as     */
/* rrtmgp_hip_interpolation_*), for inspection of the interpolation state.                    */
int rrtmgp_api_kdist_fill_col_gas(void* kdist, int ncol, int nlay, void* gas_concs,
                                  const rrtmgp_real* col_dry, rrtmgp_real* col_gas);
int rrtmgp_api_kdist_interpolation(void* kdist, int ncol, int nlay, const rrtmgp_real* play,
                                   const rrtmgp_real* tlay, const rrtmgp_real* col_gas,
                                   int* jtemp, rrtmgp_real* fmajor, rrtmgp_real* fminor,
                                   rrtmgp_real* col_mix, signed char* tropo, int* jeta,
                                   int* jpress);

/* Cloud_optics_gpu (include/Cloud_optics.h:84-128).  two_stream: 0 absorption only, 1 tau,   */
/* ssa, g, 2 tau, co-albedo, g (delta-scaled in co-albedo form when delta_scale != 0).           */
int rrtmgp_api_cloud_create(const rrtmgp_cloud_lut* lut, void** cloud);
int rrtmgp_api_cloud_destroy(void* cloud);
int rrtmgp_api_cloud_optics(void* cloud, int ncol, int nlay, const rrtmgp_real* lwp,
                            const rrtmgp_real* iwp, const rrtmgp_real* rel, const rrtmgp_real* rei,
                            int two_stream, int delta_scale, rrtmgp_real* tau, rrtmgp_real* ssa,
                            rrtmgp_real* g);

/* Aerosol_optics_gpu (include/Aerosol_optics.h:55-101): aerosol_concs is a gas-concentration */
/* handle holding aermr01..aermr11 (Aerosol_concs_gpu = Gas_concs_gpu, as in the reference);    */
/* rh (ncol, nlay), plev (ncol, nlay+1); outputs (ncol, nlay, nbnd), delta-scaled when asked,  */
/* with the co-albedo 1 - ssa in place of ssa when coalbedo != 0.                               */
int rrtmgp_api_aerosol_create(const rrtmgp_aerosol_lut* lut, void** aerosol);
int rrtmgp_api_aerosol_destroy(void* aerosol);
int rrtmgp_api_aerosol_optics(void* aerosol, int ncol, int nlay, void* aerosol_concs, const rrtmgp_real* rh,
                              const rrtmgp_real* plev, int delta_scale, int coalbedo, rrtmgp_real* tau,
                              rrtmgp_real* ssa, rrtmgp_real* g);

/* add_to and delta_scale (include/Optical_props.h:174-176,256).  nspec_in is ngpt (same grid) */
/* or nbnd (band-resolved input); ssa/g NULL for 1scl.                                        */
int rrtmgp_api_add_to(void* kdist, int ncol, int nlay, int nspec_in, rrtmgp_real* tau,
                      rrtmgp_real* ssa, rrtmgp_real* g, const rrtmgp_real* tau_in,
                      const rrtmgp_real* ssa_in, const rrtmgp_real* g_in);
/* add_to of band-resolved (ncol, nlay, nbnd) properties onto (ncol, nlay, ngpt) ones, both in    */
/* co-albedo form (rrtmgp_hip_inc_bybnd_coalbedo_*; no reference counterpart).                   */
int rrtmgp_api_add_to_coalbedo(void* kdist, int ncol, int nlay, rrtmgp_real* tau, rrtmgp_real* coalbedo,
                               rrtmgp_real* g, const rrtmgp_real* tau_in, const rrtmgp_real* coalbedo_in,
                               const rrtmgp_real* g_in);
int rrtmgp_api_delta_scale(int ncol, int nlay, int ngpt, rrtmgp_real* tau, rrtmgp_real* ssa,
                           rrtmgp_real* g);

/* Rte_lw_gpu::rte_lw / Rte_sw_gpu::rte_sw (include/Rte_lw.h:65-73, Rte_sw.h:65-75); the k-  */
/* distribution supplies the band structure.  do_broadband selects (ncol, nlev) outputs.      */
int rrtmgp_api_rte_lw(void* kdist, int ncol, int nlay, int top_at_1, const rrtmgp_real* tau,
                      const rrtmgp_real* lay_src, const rrtmgp_real* lev_src_inc,
                      const rrtmgp_real* lev_src_dec, const rrtmgp_real* sfc_src,
                      const rrtmgp_real* sfc_emis, const rrtmgp_real* inc_flux,
                      int n_gauss_angles, int do_broadband, rrtmgp_real* flux_up,
                      rrtmgp_real* flux_dn);
int rrtmgp_api_rte_sw(void* kdist, int ncol, int nlay, int top_at_1, const rrtmgp_real* tau,
                      const rrtmgp_real* ssa, const rrtmgp_real* g, const rrtmgp_real* mu0,
                      const rrtmgp_real* inc_flux_dir, const rrtmgp_real* sfc_alb_dir,
                      const rrtmgp_real* sfc_alb_dif, const rrtmgp_real* inc_flux_dif,
                      int do_broadband, rrtmgp_real* flux_up, rrtmgp_real* flux_dn,
                      rrtmgp_real* flux_dir);

/* Fluxes_broadband_gpu / Fluxes_byband_gpu::reduce (include/Fluxes.h:151-199).  gpt_dir and  */
/* the bnd_* outputs may be NULL.                                                             */
int rrtmgp_api_fluxes_reduce(void* kdist, int ncol, int nlev, const rrtmgp_real* gpt_up,
                             const rrtmgp_real* gpt_dn, const rrtmgp_real* gpt_dir,
                             rrtmgp_real* flux_up, rrtmgp_real* flux_dn, rrtmgp_real* flux_dir,
                             rrtmgp_real* flux_net, rrtmgp_real* bnd_up, rrtmgp_real* bnd_dn,
                             rrtmgp_real* bnd_dir, rrtmgp_real* bnd_net);

/* Heating rate (no reference counterpart), K/day per layer. */
int rrtmgp_api_heating_rate(int ncol, int nlay, const rrtmgp_real* flux_up,
                            const rrtmgp_real* flux_dn, const rrtmgp_real* plev, rrtmgp_real* hr);

/* Radiation_solver_{longwave,shortwave}_gpu (include_test/Radiation_solver.h:33-233).         */
/* heating_rate (ncol, nlay) may be NULL; when given it receives heating rates [K/day] formed    */
/* from each layer's net-flux convergence inside the solver, summed over g-points in FP64 (an   */
/* addition: the reference has no heating rates).                                                */
int rrtmgp_api_solver_lw_create(void* kdist, void* cloud, void** solver);
int rrtmgp_api_solver_sw_create(void* kdist, void* cloud, void** solver);
/* Shortwave solver with aerosol optics as well (aerosol may be NULL: as solver_sw_create).    */
int rrtmgp_api_solver_sw_create_aerosol(void* kdist, void* cloud, void* aerosol, void** solver);
int rrtmgp_api_solver_destroy(void* solver);
int rrtmgp_api_solver_set_block(void* solver, int n_col_block);
int rrtmgp_api_solve_lw(void* solver, int ncol, int nlay, int cloudy, int bnd_fluxes,
                        const rrtmgp_real* play, const rrtmgp_real* plev, const rrtmgp_real* tlay,
                        const rrtmgp_real* tlev, const rrtmgp_real* tsfc, void* gas_concs,
                        const rrtmgp_real* col_dry, const rrtmgp_real* emis_sfc,
                        const rrtmgp_real* lwp, const rrtmgp_real* iwp, const rrtmgp_real* rel,
                        const rrtmgp_real* rei, rrtmgp_real* flux_up, rrtmgp_real* flux_dn,
                        rrtmgp_real* flux_net, rrtmgp_real* bnd_up, rrtmgp_real* bnd_dn,
                        rrtmgp_real* bnd_net, rrtmgp_real* heating_rate);
int rrtmgp_api_solve_sw(void* solver, int ncol, int nlay, int cloudy, int bnd_fluxes,
                        int delta_cloud, const rrtmgp_real* play, const rrtmgp_real* plev,
                        const rrtmgp_real* tlay, const rrtmgp_real* tlev, void* gas_concs,
                        const rrtmgp_real* col_dry, const rrtmgp_real* sfc_alb_dir,
                        const rrtmgp_real* sfc_alb_dif, const rrtmgp_real* tsi_scaling,
                        const rrtmgp_real* mu0, const rrtmgp_real* lwp, const rrtmgp_real* iwp,
                        const rrtmgp_real* rel, const rrtmgp_real* rei, rrtmgp_real* flux_up,
                        rrtmgp_real* flux_dn, rrtmgp_real* flux_dir, rrtmgp_real* flux_net,
                        rrtmgp_real* bnd_up, rrtmgp_real* bnd_dn, rrtmgp_real* bnd_dir,
                        rrtmgp_real* bnd_net, rrtmgp_real* heating_rate);

/* rrtmgp_api_solve_sw with aerosols (src_test/Radiation_solver.cu:794-810): aerosol != 0      */
/* adds the aerosol optical properties of rh (ncol, nlay) and aerosol_concs (a gas-concentration */
/* handle of aermr01..aermr11) after the clouds, delta-scaled when delta_aerosol != 0.          */
int rrtmgp_api_solve_sw_aerosol(void* solver, int ncol, int nlay, int cloudy, int aerosol, int bnd_fluxes,
                                int delta_cloud, int delta_aerosol, const rrtmgp_real* play,
                                const rrtmgp_real* plev, const rrtmgp_real* tlay, const rrtmgp_real* tlev,
                                void* gas_concs, const rrtmgp_real* col_dry, const rrtmgp_real* sfc_alb_dir,
                                const rrtmgp_real* sfc_alb_dif, const rrtmgp_real* tsi_scaling,
                                const rrtmgp_real* mu0, const rrtmgp_real* lwp, const rrtmgp_real* iwp,
                                const rrtmgp_real* rel, const rrtmgp_real* rei, const rrtmgp_real* rh,
                                void* aerosol_concs, rrtmgp_real* flux_up, rrtmgp_real* flux_dn,
                                rrtmgp_real* flux_dir, rrtmgp_real* flux_net, rrtmgp_real* bnd_up,
                                rrtmgp_real* bnd_dn, rrtmgp_real* bnd_dir, rrtmgp_real* bnd_net,
                                rrtmgp_real* heating_rate);

/* Adjoints of the two drivers above (no reference counterpart): for cotangents g_flux_* of   */
/* the broadband fluxes (ncol, nlev), the gradients of L = sum(g_flux * flux) with respect to */
/* tlay, tlev, tsfc (LW), ln play, plev (through col_dry, which is computed from h2o), the    */
/* vmr of every k-distribution gas (d_vmr (ncol, nlay, ngas), k-distribution gas order) and   */
/* the surface emissivity / albedos ((nbnd, ncol)).  Clouds (cloudy != 0) are held fixed.     */
/* Gas optics is recomputed band by band.  All arrays are device arrays; outputs overwritten. */
int rrtmgp_api_adjoint_lw(void* solver, int ncol, int nlay, int cloudy, const rrtmgp_real* play,
                          const rrtmgp_real* plev, const rrtmgp_real* tlay, const rrtmgp_real* tlev,
                          const rrtmgp_real* tsfc, void* gas_concs, const rrtmgp_real* emis_sfc,
                          const rrtmgp_real* lwp, const rrtmgp_real* iwp, const rrtmgp_real* rel,
                          const rrtmgp_real* rei, const rrtmgp_real* g_flux_up,
                          const rrtmgp_real* g_flux_dn, rrtmgp_real* d_tlay, rrtmgp_real* d_tlev,
                          rrtmgp_real* d_tsfc, rrtmgp_real* d_lnp_lay, rrtmgp_real* d_plev,
                          rrtmgp_real* d_vmr, rrtmgp_real* d_emis_sfc);
int rrtmgp_api_adjoint_sw(void* solver, int ncol, int nlay, int cloudy, int delta_cloud,
                          const rrtmgp_real* play, const rrtmgp_real* plev, const rrtmgp_real* tlay,
                          void* gas_concs, const rrtmgp_real* sfc_alb_dir,
                          const rrtmgp_real* sfc_alb_dif, const rrtmgp_real* tsi_scaling,
                          const rrtmgp_real* mu0, const rrtmgp_real* lwp, const rrtmgp_real* iwp,
                          const rrtmgp_real* rel, const rrtmgp_real* rei,
                          const rrtmgp_real* g_flux_up, const rrtmgp_real* g_flux_dn,
                          const rrtmgp_real* g_flux_dir, rrtmgp_real* d_tlay,
                          rrtmgp_real* d_lnp_lay, rrtmgp_real* d_plev, rrtmgp_real* d_vmr,
                          rrtmgp_real* d_sfc_alb_dir, rrtmgp_real* d_sfc_alb_dif);

#ifdef __cplusplus
}
#endif
#endif
