// Longwave source holder (reference src_cuda/Source_functions.cu:30-40).
#include "Source_functions.h"

Source_func_lw_gpu::Source_func_lw_gpu(const int n_col, const int n_lay, const Optical_props_gpu& p) :
    Optical_props_gpu(p),
    sfc_source({n_col, p.get_ngpt()}),
    sfc_source_jac({n_col, p.get_ngpt()}),
    lay_source({n_col, n_lay, p.get_ngpt()}),
    lev_source_inc({n_col, n_lay, p.get_ngpt()}),
    lev_source_dec({n_col, n_lay, p.get_ngpt()})
{}

Source_func_lw_gpu::Source_func_lw_gpu(const int n_col, const int n_lay, const Optical_props_gpu& p,
                                       bool factored_) :
    Optical_props_gpu(p),
    sfc_source({n_col, p.get_ngpt()}),
    sfc_source_jac({n_col, p.get_ngpt()}),
    lay_source(factored_ ? Array_gpu<Float,3>() : Array_gpu<Float,3>({n_col, n_lay, p.get_ngpt()})),
    lev_source_inc(factored_ ? Array_gpu<Float,3>() : Array_gpu<Float,3>({n_col, n_lay, p.get_ngpt()})),
    lev_source_dec(factored_ ? Array_gpu<Float,3>() : Array_gpu<Float,3>({n_col, n_lay, p.get_ngpt()})),
    factored(factored_),
    planck_frac(factored_ ? Array_gpu<Float,3>({n_col, n_lay, p.get_ngpt()}) : Array_gpu<Float,3>()),
    band_lay_source(factored_ ? Array_gpu<Float,3>({n_col, n_lay, p.get_nband()}) : Array_gpu<Float,3>()),
    band_lev_source(factored_ ? Array_gpu<Float,3>({n_col, n_lay + 1, p.get_nband()}) : Array_gpu<Float,3>())
{}
