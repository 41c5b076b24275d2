// Optical-property classes (reference src_cuda/Optical_props.cu:33-165, src/Optical_props.cpp).
#include "Optical_props.h"

#include "hip_dispatch.h"

namespace rrtmgp_cpp
{
namespace { thread_local void* g_stream = nullptr; }
void* stream() { return g_stream; }
void set_stream(void* s) { g_stream = s; }
}

using rrtmgp_cpp::hip_check;
using rrtmgp_cpp::stream;

Optical_props_gpu::Optical_props_gpu(const Array<Float,2>& band_lims_wvn, const Array<int,2>& band_lims_gpt) :
    band2gpt(band_lims_gpt), band_lims_wvn(band_lims_wvn)
{
    init_maps();
}

// Band-resolved properties: one "g-point" per band (src/Optical_props.cpp:51-72).
Optical_props_gpu::Optical_props_gpu(const Array<Float,2>& band_lims_wvn) : band_lims_wvn(band_lims_wvn)
{
    Array<int,2> lims({2, band_lims_wvn.dim(2)});
    for (int b = 1; b <= band_lims_wvn.dim(2); ++b) { lims({1, b}) = b; lims({2, b}) = b; }
    band2gpt = lims;
    init_maps();
}

void Optical_props_gpu::init_maps()
{
    gpt2band = Array<int,1>({band2gpt.max()});
    for (int b = 1; b <= band2gpt.dim(2); ++b)
        for (int g = band2gpt({1, b}); g <= band2gpt({2, b}); ++g)
            gpt2band({g}) = b;
    band2gpt_gpu = band2gpt;
    gpt2band_gpu = gpt2band;
}

Optical_props_1scl_gpu::Optical_props_1scl_gpu(const int ncol, const int nlay, const Optical_props_gpu& p) :
    Optical_props_arry_gpu(p), tau({ncol, nlay, p.get_ngpt()})
{}

Optical_props_2str_gpu::Optical_props_2str_gpu(const int ncol, const int nlay, const Optical_props_gpu& p) :
    Optical_props_arry_gpu(p), tau({ncol, nlay, p.get_ngpt()}), ssa({ncol, nlay, p.get_ngpt()}),
    g({ncol, nlay, p.get_ngpt()})
{}

// delta_scale (src/Optical_props.cpp:207-216).
void Optical_props_2str_gpu::delta_scale(const Array_gpu<Float,3>&)
{
    hip_check(K<Float>::delta_scale_2str(get_ncol(), get_nlay(), get_ngpt(), tau.ptr(), ssa.ptr(), g.ptr(), stream()));
}

// add_to (src/Optical_props.cpp:219-268): same spectral grid, or band-resolved props added to
// every g-point of their band.
void add_to(Optical_props_1scl_gpu& a, const Optical_props_1scl_gpu& b)
{
    const int ncol = a.get_ncol(), nlay = a.get_nlay(), ngpt = a.get_ngpt();
    if (ngpt == b.get_ngpt())
        hip_check(K<Float>::increment(ncol, nlay, ngpt, a.get_tau().ptr(), nullptr, nullptr,
                                      b.get_tau().ptr(), nullptr, nullptr, stream()));
    else
    {
        if (b.get_ngpt() != a.get_nband())
            throw std::runtime_error("Cannot add optical properties with incompatible band - gpoint combination");
        hip_check(K<Float>::inc_bybnd(ncol, nlay, ngpt, a.get_tau().ptr(), nullptr, nullptr, b.get_tau().ptr(),
                                      nullptr, nullptr, a.get_nband(), a.get_band_lims_gpoint_gpu().ptr(), stream()));
    }
}

void add_to(Optical_props_2str_gpu& a, const Optical_props_2str_gpu& b)
{
    const int ncol = a.get_ncol(), nlay = a.get_nlay(), ngpt = a.get_ngpt();
    if (ngpt == b.get_ngpt())
        hip_check(K<Float>::increment(ncol, nlay, ngpt, a.get_tau().ptr(), a.get_ssa().ptr(), a.get_g().ptr(),
                                      b.get_tau().ptr(), b.get_ssa().ptr(), b.get_g().ptr(), stream()));
    else
    {
        if (b.get_ngpt() != a.get_nband())
            throw std::runtime_error("Cannot add optical properties with incompatible band - gpoint combination");
        hip_check(K<Float>::inc_bybnd(ncol, nlay, ngpt, a.get_tau().ptr(), a.get_ssa().ptr(), a.get_g().ptr(),
                                      b.get_tau().ptr(), b.get_ssa().ptr(), b.get_g().ptr(), a.get_nband(),
                                      a.get_band_lims_gpoint_gpu().ptr(), stream()));
    }
}

void add_to_coalbedo(Optical_props_2str_gpu& a, const Optical_props_2str_gpu& b)
{
    if (b.get_ngpt() != a.get_nband())
        throw std::runtime_error("add_to_coalbedo: the input must be band-resolved on the output's bands");
    hip_check(K<Float>::inc_bybnd_coalbedo(a.get_ncol(), a.get_nlay(), a.get_ngpt(), a.get_tau().ptr(),
                                           a.get_ssa().ptr(), a.get_g().ptr(), b.get_tau().ptr(), b.get_ssa().ptr(),
                                           b.get_g().ptr(), a.get_nband(), a.get_band_lims_gpoint_gpu().ptr(),
                                           stream()));
}
