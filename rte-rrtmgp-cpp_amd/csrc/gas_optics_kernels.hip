// Gas optics kernels for gfx950.
//
// Arithmetic follows the reference restatement in src_kernels_cuda/gas_optics_rrtmgp_kernels.cu
// (interpolation :215-293, major :296-341, minor :356-476, Rayleigh :572-616, Planck :129-212,
// interpolate1D :1-12) with the CPU combine rule of src/Gas_optics_rrtmgp.cpp:365-385.
//
// Design.  In the fused kernels a work-group (4 wavefronts) computes one layer of 64 columns
// for one chunk of at most 16 g-points of a band, in three phases:
//   1. wavefront 0, lane = column: the pair's interpolation state for the band's flavor
//      (indices, weights, minor scalings, Planck values) -> LDS;
//   2. lane = (pair, quad of 4 g-points): every lane reads one table corner of its four
//      g-points with one 16-byte load from band-packed tables (g-points of a band fastest), so
//      the 4 lanes of a pair read one contiguous 64-byte segment; results -> LDS;
//   3. lane = column, wavefront = g-point row: finishing arithmetic and 256-byte coalesced
//      stores in the reference layout (ncol, nlay, ngpt).
// LDS rows are padded to 68 floats (conflict-free for both lane layouts).  Work-groups of one
// chunk are consecutive and kept on one XCD, so that XCD's L2 holds the chunk's table slice.
#include <algorithm>
#include <memory>
#include <vector>

#include "common.h"

namespace rrtmgp_hip
{
constexpr int GO_THREADS = 256;   // work-group: 4 wavefronts
constexpr int GO_COLS = 64;       // columns of a work-group's layer
constexpr int GO_GPTS = 16;       // g-points per chunk: the lanes of one pair
constexpr int GO_ROW = 68;        // LDS row length (64 columns + padding)

// Unpadded (g-point row, column) LDS tiles of the forward kernels: column c of row r sits at
// c ^ 16 * (r / 4 % 4), so phase 2's stores (four rows gq + e, 16 columns each, per instruction)
// and phase 3's row reads fall on distinct banks.
__host__ __device__ constexpr int go_swz(int r, int c) { return c ^ (((r >> 2) & 3) << 4); }
constexpr int GO_MAXM = 8;        // minor intervals per band and half of the atmosphere

// One minor interval of a band (one half of the atmosphere), as gas indices of the col_gas
// numbering (0 = dry air).
struct GoMinor
{
    int gas;          // idx_minor
    int dens;         // minor_scales_with_density
    int scal_gas;     // idx_minor_scaling (<= 0: none)
    int compl_;       // scale_by_complement
};

// Descriptor of one band for the fused kernels; half 0 is the lower atmosphere (tropo), 1 the
// upper.  Offsets are in elements of the band-packed tables.
struct GoBand
{
    int g0, ng, gp;          // first g-point (0-based), g-points, g-points padded to a multiple of 4
    int koff, roff, moff;    // offsets in kmajor_p / planck_frac_p, krayl_p, kminor_p
    int nslot;               // minor slots per half: max(nq[0], nq[1])
    int nq[2];               // minor intervals of the band per half
    int gas[2][2];           // the band's flavor gases per half
    GoMinor minor[2][GO_MAXM];
};

template<typename Real>
struct Kd
{
    int ngas, nflav, neta, npres, ntemp, ngpt, nbnd, nplanck, idx_h2o;
    Real press_ref_log_delta, temp_ref_min, temp_ref_delta, press_ref_trop_log, totplnk_delta;
    int nminor[2];                 // [0] lower (tropo), [1] upper: indexed by itropo = !tropo
    const int* flavor;
    const int* gpoint_flavor;
    const int* band_lims_gpt;
    const Real* press_ref_log;
    const Real* temp_ref;
    const Real* vmr_ref;
    const Real* kmajor;
    const Real* kminor[2];
    const int* minor_limits[2];
    const Bool* dens[2];
    const Bool* compl_[2];
    const int* idx_minor[2];
    const int* idx_scal[2];
    const int* kstart[2];
    const Real* totplnk;
    const Real* planck_frac;
    const Real* krayl;
    // Band-packed tables of the fused kernels (kdist_pack)
    const struct GoBand* bands;
    const Real* kmajor_p;
    const Real* planck_frac_p;
    const Real* kminor_p;
    const Real* krayl_p;
    int max_band_gpts;
};

template<typename Real>
Kd<Real> make_kd(const rrtmgp_kdist_dev& d)
{
    Kd<Real> k;
    k.ngas = int(d.ngas); k.nflav = int(d.nflav); k.neta = int(d.neta); k.npres = int(d.npres);
    k.ntemp = int(d.ntemp); k.ngpt = int(d.ngpt); k.nbnd = int(d.nbnd); k.nplanck = int(d.nplanck);
    k.idx_h2o = int(d.idx_h2o);
    k.press_ref_log_delta = Real(d.press_ref_log_delta);
    k.temp_ref_min = Real(d.temp_ref_min);
    k.temp_ref_delta = Real(d.temp_ref_delta);
    k.press_ref_trop_log = Real(d.press_ref_trop_log);
    k.totplnk_delta = Real(d.totplnk_delta);
    k.nminor[0] = int(d.nminor_lower); k.nminor[1] = int(d.nminor_upper);
    k.flavor = d.flavor; k.gpoint_flavor = d.gpoint_flavor; k.band_lims_gpt = d.band_lims_gpt;
    k.press_ref_log = static_cast<const Real*>(d.press_ref_log);
    k.temp_ref = static_cast<const Real*>(d.temp_ref);
    k.vmr_ref = static_cast<const Real*>(d.vmr_ref);
    k.kmajor = static_cast<const Real*>(d.kmajor);
    k.kminor[0] = static_cast<const Real*>(d.kminor_lower);
    k.kminor[1] = static_cast<const Real*>(d.kminor_upper);
    k.minor_limits[0] = d.minor_limits_gpt_lower; k.minor_limits[1] = d.minor_limits_gpt_upper;
    k.dens[0] = d.minor_scales_with_density_lower; k.dens[1] = d.minor_scales_with_density_upper;
    k.compl_[0] = d.scale_by_complement_lower; k.compl_[1] = d.scale_by_complement_upper;
    k.idx_minor[0] = d.idx_minor_lower; k.idx_minor[1] = d.idx_minor_upper;
    k.idx_scal[0] = d.idx_minor_scaling_lower; k.idx_scal[1] = d.idx_minor_scaling_upper;
    k.kstart[0] = d.kminor_start_lower; k.kstart[1] = d.kminor_start_upper;
    k.totplnk = static_cast<const Real*>(d.totplnk);
    k.planck_frac = static_cast<const Real*>(d.planck_frac);
    k.krayl = static_cast<const Real*>(d.krayl);
    k.bands = static_cast<const GoBand*>(d.go_bands);
    k.kmajor_p = static_cast<const Real*>(d.kmajor_p);
    k.planck_frac_p = static_cast<const Real*>(d.planck_frac_p);
    k.kminor_p = static_cast<const Real*>(d.kminor_p);
    k.krayl_p = static_cast<const Real*>(d.krayl_p);
    k.max_band_gpts = int(d.max_band_gpts);
    return k;
}

template<typename Real>
struct Gases
{
    const Real* p[RRTMGP_MAX_GAS];
    int64_t scol[RRTMGP_MAX_GAS];
    int64_t slay[RRTMGP_MAX_GAS];
};

template<typename Real>
int make_gases(const rrtmgp_gas_vmr* g, int64_t ngas_kd, int64_t ncol, int64_t nlay, Gases<Real>& out)
{
    if (!g) return arg_error("gas concentrations are NULL");
    if (g->ngas != ngas_kd)
        return arg_error("gas concentrations hold " + std::to_string(g->ngas) +
                         " gases, the k-distribution expects " + std::to_string(ngas_kd));
    for (int i = 0; i < g->ngas; ++i)
    {
        const int64_t d1 = g->dim1[i], d2 = g->dim2[i];
        if (!g->vmr[i]) return arg_error("vmr pointer of gas " + std::to_string(i + 1) + " is NULL");
        if (!((d1 == 1 || d1 == ncol) && (d2 == 1 || d2 == nlay)) || (d1 == ncol && ncol > 1 && d2 == 1))
            return arg_error("vmr of gas " + std::to_string(i + 1) + " has dims (" + std::to_string(d1) +
                             ", " + std::to_string(d2) + "); expected (1|ncol, 1|nlay)");
        out.p[i] = static_cast<const Real*>(g->vmr[i]);
        out.scol[i] = (d1 == 1) ? 0 : 1;
        out.slay[i] = (d2 == 1) ? 0 : d1;
    }
    return RRTMGP_OK;
}

// col_gas of gas index `gas` (0 = dry air): vmr * col_dry (compute_gas_taus, :1150-1154).
template<typename Real>
__device__ __forceinline__ Real col_of(const Gases<Real>& g, int gas, int64_t icol, int64_t ilay, Real col_dry)
{
    if (gas == 0) return col_dry;
    const int i = gas - 1;
    return g.p[i][icol * g.scol[i] + ilay * g.slay[i]] * col_dry;
}

template<typename Real>
struct CL { int jtemp, jpress, itropo; Real ftemp, fpress; };

template<typename Real>
struct FL { int jeta[2]; Real col_mix[2]; Real fminor[2][2]; Real fmajor[2][4]; };

// Per (col, lay) part of interpolation_kernel (.cu:236-249).  FMA contraction is off so the
// FP64 indices and weights reproduce a contraction-free CPU evaluation bit for bit.
template<typename Real>
__device__ __forceinline__ CL<Real> interp_cl(const Kd<Real>& k, Real p, Real t)
{
#pragma clang fp contract(off)
    CL<Real> c;
    int jt = int((t - (k.temp_ref_min - k.temp_ref_delta)) / k.temp_ref_delta);
    jt = min(k.ntemp - 1, max(1, jt));
    c.jtemp = jt;
    c.ftemp = (t - k.temp_ref[jt - 1]) / k.temp_ref_delta;
    const Real lp = log(p);
    const Real locpress = Real(1) + (lp - k.press_ref_log[0]) / k.press_ref_log_delta;
    c.jpress = min(k.npres - 1, max(1, int(locpress)));
    c.fpress = locpress - Real(c.jpress);
    c.itropo = (lp > k.press_ref_trop_log) ? 0 : 1;
    return c;
}

// fmod(x, 1) without the general remainder routine: x - trunc(x) is exact in floating point (the
// difference of x and its integer part is representable), and taking x's sign keeps fmod's signed
// zero for negative integers; infinities and NaN give NaN as fmod does.
template<typename Real>
__device__ __forceinline__ Real frac1(Real x) { return copysign(x - trunc(x), x); }

// Per-flavor part of interpolation_kernel (.cu:254-291).
template<typename Real>
__device__ __forceinline__ void interp_flav(const Kd<Real>& k, const CL<Real>& c, int gas1, int gas2,
                                            Real col1, Real col2, FL<Real>& f)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int itemp = 0; itemp < 2; ++itemp)
    {
        const int64_t vb = c.itropo + int64_t(c.jtemp + itemp - 1) * (k.ngas + 1) * 2;
        const Real ratio = k.vmr_ref[vb + 2 * gas1] / k.vmr_ref[vb + 2 * gas2];
        const Real cm = col1 + ratio * col2;
        f.col_mix[itemp] = cm;
        const Real eta = (cm > Real(2) * Limits<Real>::tiny) ? col1 / cm : Real(0.5);
        const Real loceta = eta * Real(k.neta - 1);
        // The lower clamp only matters for non-finite inputs; it keeps every gather in bounds.
        f.jeta[itemp] = max(1, min(int(loceta) + 1, k.neta - 1));
        const Real feta = frac1(loceta);
        const Real ftemp_term = Real(1 - itemp) + Real(2 * itemp - 1) * c.ftemp;
        const Real m0 = (Real(1) - feta) * ftemp_term;
        const Real m1 = feta * ftemp_term;
        f.fminor[itemp][0] = m0;
        f.fminor[itemp][1] = m1;
        f.fmajor[itemp][0] = (Real(1) - c.fpress) * m0;
        f.fmajor[itemp][1] = (Real(1) - c.fpress) * m1;
        f.fmajor[itemp][2] = c.fpress * m0;
        f.fmajor[itemp][3] = c.fpress * m1;
    }
}

// Major absorption of one g-point (.cu:317-339); kmajor is (ntemp, neta, npres+1, ngpt).
template<typename Real>
__device__ __forceinline__ Real major_tau(const Kd<Real>& k, const CL<Real>& c, const FL<Real>& f, int igpt)
{
    const int jpi = c.jpress + c.itropo;
    const int64_t se = k.ntemp, sp = int64_t(k.ntemp) * k.neta;
    Real tau = Real(0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        const int64_t b = (c.jtemp - 1 + i) + se * (f.jeta[i] - 1) + sp * ((jpi - 1) + int64_t(k.npres + 1) * igpt);
        const Real* km = k.kmajor + b;
        tau += f.col_mix[i] * (f.fmajor[i][0] * km[0] + f.fmajor[i][1] * km[se] +
                               f.fmajor[i][2] * km[sp] + f.fmajor[i][3] * km[sp + se]);
    }
    return tau;
}

// Planck fraction of one g-point (.cu:170-180); planck_frac has the kmajor layout.
template<typename Real>
__device__ __forceinline__ Real planck_fraction(const Kd<Real>& k, const CL<Real>& c, const FL<Real>& f, int igpt)
{
    const int jp = c.jpress + c.itropo;
    const int64_t se = k.ntemp, sp = int64_t(k.ntemp) * k.neta;
    const Real* p0 = k.planck_frac + (c.jtemp - 1) + se * (f.jeta[0] - 1) + sp * ((jp - 1) + int64_t(k.npres + 1) * igpt);
    const Real* p1 = k.planck_frac + (c.jtemp) + se * (f.jeta[1] - 1) + sp * ((jp - 1) + int64_t(k.npres + 1) * igpt);
    return (f.fmajor[0][0] * p0[0] + f.fmajor[0][1] * p0[se] + f.fmajor[0][2] * p0[sp] + f.fmajor[0][3] * p0[sp + se])
         + (f.fmajor[1][0] * p1[0] + f.fmajor[1][1] * p1[se] + f.fmajor[1][2] * p1[sp] + f.fmajor[1][3] * p1[sp + se]);
}

// interpolate1D (.cu:1-12), in two steps: the position of val in the table (index and fraction),
// which the fused kernels compute once per column for all bands, and the lookup.
template<typename Real>
__device__ __forceinline__ void interp1d_pos(Real val, Real offset, Real delta, int len, int& idx, Real& frac)
{
    const Real val0 = (val - offset) / delta;
    frac = val0 - Real(int(val0));
    idx = min(len - 1, max(1, int(val0) + 1));
}

template<typename Real>
__device__ __forceinline__ Real interp1d_at(const Real* table, int idx, Real frac)
{
    return table[idx - 1] + frac * (table[idx] - table[idx - 1]);
}

template<typename Real>
__device__ __forceinline__ Real interpolate1D(Real val, Real offset, Real delta, int len, const Real* table)
{
    int idx;
    Real frac;
    interp1d_pos(val, offset, delta, len, idx, frac);
    return interp1d_at(table, idx, frac);
}

// Scaling of one minor interval at one (col, lay) (.cu:388-420); col(gas) gives the column
// amount of a gas index (0 = dry air).
template<typename Real, typename ColOf>
__device__ __forceinline__ Real minor_scaling_of(const ColOf& col, int gas, bool dens, int scal_gas, bool compl_,
                                                 int idx_h2o, Real p, Real t, Real col_dry)
{
    Real scaling = col(gas);
    if (dens)
    {
        const Real PaTohPa = Real(0.01);
        scaling *= PaTohPa * p / t;
        if (scal_gas > 0)
        {
            const Real vmr_fact = Real(1) / col_dry;
            const Real dry_fact = Real(1) / (Real(1) + col(idx_h2o) * vmr_fact);
            const Real cs = col(scal_gas);
            if (compl_)
                scaling *= (Real(1) - cs * vmr_fact * dry_fact);
            else
                scaling *= cs * vmr_fact * dry_fact;
        }
    }
    return scaling;
}

// ---------------------------------------------------------------------------------------------
// Fused kernels.
//
// They read the band-packed tables of kdist_pack.  Per band the band's g-points (padded with
// zeros to gp, a multiple of 4) are the fastest index, so one 16-byte read yields one table
// corner for four consecutive g-points:
//   kmajor_p, planck_frac_p  (gp, npres+1, neta, ntemp)        at GoBand::koff
//   krayl_p                  (gp, neta, ntemp, half)            at GoBand::roff
//   kminor_p                 (gp, nslot, neta, ntemp, half)     at GoBand::moff
// Slot q of half s of kminor_p holds the band's q-th minor interval of that half (ascending
// interval order) and is zero at the g-points outside the interval: the minor sum runs over the
// band's slots with no per-g-point interval test, and a g-point outside an interval adds an
// exact zero, so the sum equals the reference's sum over the intervals containing the g-point.

// Grid of the fused kernels: one work-group per (64-column tile, layer) unit, tile fastest; the
// work-group runs the launch's chunks (band, 16-g-point part of the band) one after the other.
// A launch covers bands b0 .. b0+nbnd-1; its outputs hold their g-points only, counted from the
// first g-point of band b0.  With a tile list the units are tiles[0 .. nlist-1] (each
// ilay * ncoltiles + tile) instead of all of them.
struct GoGrid
{
    int64_t ncoltiles, nlay, cpb;
    int b0, nbnd;
    const int* tiles;
    int64_t nlist;
    int64_t ntl;           // (tile, layer) units: the work-groups of the forward kernels
};

inline GoGrid make_go_grid(int64_t ncol, int64_t nlay, int64_t b0, int64_t nbnd, int64_t max_band_gpts,
                           const int* tiles = nullptr, int64_t nlist = 0)
{
    GoGrid t;
    t.ncoltiles = (ncol + GO_COLS - 1) / GO_COLS;
    t.nlay = nlay;
    t.cpb = std::max<int64_t>(1, (max_band_gpts + GO_GPTS - 1) / GO_GPTS);
    t.tiles = tiles;
    t.nlist = tiles ? nlist : 0;
    t.ntl = tiles ? nlist : t.ncoltiles * nlay;
    t.b0 = int(b0);
    t.nbnd = int(nbnd);
    return t;
}

// (64-column tile, layer) unit u of the forward kernels: tile fastest, or the listed ids
// (ilay * ncoltiles + tile).  ok is false for a listed id outside the grid.
struct GoTl
{
    int64_t ilay, col0;
    bool ok;
};

__device__ __forceinline__ GoTl go_tl(const GoGrid& t, int64_t u64)
{
    GoTl r;
    const uint32_t u = uint32_t(u64), nct = uint32_t(t.ncoltiles);
    const uint32_t id = t.tiles ? uint32_t(t.tiles[u]) : u;
    const uint32_t ilay = id / nct;
    r.ilay = ilay;
    r.col0 = int64_t(id - ilay * nct) * GO_COLS;
    r.ok = ilay < uint32_t(t.nlay);
    return r;
}

// Volume-mixing-ratio pointers and strides, staged in LDS so that a per-lane gas index reads
// them with an ordinary LDS load.
template<typename Real>
struct GasRef
{
    const Real* p;
    int64_t scol, slay;
};

template<typename Real>
__device__ __forceinline__ void stage_gases(const Gases<Real>& g, int ngas, GasRef<Real>* sg)
{
    for (int i = 0; i < ngas; ++i)
        sg[i] = GasRef<Real>{g.p[i], g.scol[i], g.slay[i]};
}

template<typename Real>
__device__ __forceinline__ Real col_lds(const GasRef<Real>* sg, int gas, int64_t icol, int64_t ilay, Real col_dry)
{
    if (gas == 0) return col_dry;
    const GasRef<Real>& r = sg[gas - 1];
    return r.p[icol * r.scol + ilay * r.slay] * col_dry;
}

// State of one (column, layer) pair for one band, computed once per pair (phase 1 of the fused
// kernels) and read from LDS by the four lanes that share the pair (phase 2).
template<typename Real>
struct PairState
{
    int om[2];           // kmajor_p / planck_frac_p offsets in the band for itemp 0/1, no g-point
    int mn[2];           // kminor_p offsets in the band (half included) for itemp 0/1
    int ry[2];           // krayl_p offsets in the band (half included) for itemp 0/1
    int flags;           // bit 0: valid column
    Real col_mix[2];
    Real fmajor[2][4];
    Real fminor[2][2];
    Real scaling[GO_MAXM];
    Real aux[5];         // LW: B(T_lay), B(T_lev above), B(T_lev below), B(T_sfc), B(T_sfc + 1); SW: col_ray
};

// The per-(column, layer) inputs of phase 1, read from memory ...
template<typename Real>
struct PairInGlobal
{
    const GasRef<Real>* sg;
    int64_t icol, ilay, ncol;
    const Real *play, *tlay, *tlev, *tsfc, *col_dry;
    __device__ Real p() const { return play[icol + ncol * ilay]; }
    __device__ Real t() const { return tlay[icol + ncol * ilay]; }
    __device__ Real cd() const { return col_dry[icol + ncol * ilay]; }
    __device__ Real col(int gas, Real cdv) const { return col_lds(sg, gas, icol, ilay, cdv); }
    __device__ Real t_hi() const { return tlev[icol + ncol * (ilay + 1)]; }
    __device__ Real t_lo() const { return tlev[icol + ncol * ilay]; }
    __device__ Real t_sfc() const { return tsfc[icol]; }
};

// ... or from the work-group's LDS rows, filled by all its threads at once (one memory round trip
// instead of the dependent loads of one wavefront) and read by every chunk of the work-group:
// row 0 p, 1 T, 2 col_dry, 3 .. 3 + ngas - 1 the vmr of gases 1 .. ngas, then (longwave) T of the
// level below and above and T_sfc.
constexpr int GO_IN_FIXED = 3;
constexpr int GO_IN_MAXROWS = GO_IN_FIXED + RRTMGP_MAX_GAS + 3;   // the most input rows of a launch

template<typename Real>
struct PairInLds
{
    const Real* rows;
    int lane, ngas;
    __device__ Real v(int r) const { return rows[r * GO_COLS + lane]; }
    __device__ Real p() const { return v(0); }
    __device__ Real t() const { return v(1); }
    __device__ Real cd() const { return v(2); }
    __device__ Real col(int gas, Real cdv) const { return gas == 0 ? cdv : v(GO_IN_FIXED + gas - 1) * cdv; }
    __device__ Real t_lo() const { return v(GO_IN_FIXED + ngas); }
    __device__ Real t_hi() const { return v(GO_IN_FIXED + ngas + 1); }
    __device__ Real t_sfc() const { return v(GO_IN_FIXED + ngas + 2); }
};

template<bool LW>
__host__ __device__ constexpr int go_in_rows(int ngas) { return GO_IN_FIXED + ngas + (LW ? 3 : 0); }

// Fills the LDS input rows of the work-group's 64 columns of layer ilay (all threads; the
// caller synchronises before reading them).
template<typename Real, bool LW>
__device__ __forceinline__ void stage_pair_inputs(const GasRef<Real>* sg, int ngas, int64_t col0, int64_t ilay,
                                                  int64_t ncol, int64_t nlay, const Real* play, const Real* tlay,
                                                  const Real* tlev, const Real* tsfc, const Real* col_dry,
                                                  Real* rows)
{
    // All loads first (row r is uniform over a wavefront), then the LDS stores.
    constexpr int NIT = (GO_IN_MAXROWS * GO_COLS + GO_THREADS - 1) / GO_THREADS;
    const int n = go_in_rows<LW>(ngas) * GO_COLS;
    Real v[NIT];
#pragma unroll
    for (int i = 0; i < NIT; ++i)
    {
        const int e = int(threadIdx.x) + i * GO_THREADS;
        v[i] = Real(0);
        if (e >= n) continue;
        const int r = e / GO_COLS, c = e % GO_COLS;
        const int64_t icol = min(col0 + c, ncol - 1);           // padding lanes repeat the last column
        const int64_t idx = icol + ncol * ilay;
        if (r == 0) v[i] = play[idx];
        else if (r == 1) v[i] = tlay[idx];
        else if (r == 2) v[i] = col_dry[idx];
        else if (r < GO_IN_FIXED + ngas)
        {
            const GasRef<Real>& g = sg[r - GO_IN_FIXED];
            v[i] = g.p[icol * g.scol + ilay * g.slay];
        }
        else if (r == GO_IN_FIXED + ngas) v[i] = tlev[idx];
        else if (r == GO_IN_FIXED + ngas + 1) v[i] = tlev[icol + ncol * (ilay + 1)];
        else v[i] = tsfc[icol];
    }
#pragma unroll
    for (int i = 0; i < NIT; ++i)
    {
        const int e = int(threadIdx.x) + i * GO_THREADS;
        if (e < n) rows[e] = v[i];
    }
}

// Phase 1 for one pair (lane = column), in parts: the interpolation state of the major gathers
// (ps_major), the minor scalings of slots q0, q0 + dq, ... (ps_minor), and the band Planck values
// (LW) or the Rayleigh column (SW) (ps_aux).  pair_state_in runs them all; the fused forward
// kernels share them out over their four wavefronts (pair_state_split).
template<typename Real, typename In>
__device__ __forceinline__ void ps_major(const Kd<Real>& k, const GoBand& bd, const CL<Real>& c, const In& in, Real cd,
                                         PairState<Real>& st)
{
    const int s = c.itropo;
    const int gas1 = bd.gas[s][0], gas2 = bd.gas[s][1];
    FL<Real> f;
    interp_flav(k, c, gas1, gas2, in.col(gas1, cd), in.col(gas2, cd), f);
    const int gp = bd.gp, np1 = k.npres + 1, nte = k.ntemp * k.neta;
    const int jp = c.jpress + c.itropo - 1;
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        const int te = (c.jtemp - 1 + i) * k.neta + (f.jeta[i] - 1);     // (temperature, eta) node
        st.om[i] = (te * np1 + jp) * gp;
        st.mn[i] = (s * nte + te) * bd.nslot * gp;
        st.ry[i] = (s * nte + te) * gp;
        st.col_mix[i] = f.col_mix[i];
#pragma unroll
        for (int q = 0; q < 4; ++q) st.fmajor[i][q] = f.fmajor[i][q];
        st.fminor[i][0] = f.fminor[i][0];
        st.fminor[i][1] = f.fminor[i][1];
    }
    st.flags = 1;
}

template<typename Real, typename In>
__device__ __forceinline__ void ps_minor(const Kd<Real>& k, const GoBand& bd, int s, const In& in, Real pv, Real tv,
                                         Real cd, int q0, int dq, PairState<Real>& st)
{
    const int nq = bd.nq[s];
    for (int q = q0; q < bd.nslot; q += dq)
    {
        Real sc = Real(0);
        if (q < nq)
        {
            const GoMinor md = bd.minor[s][q];
            sc = minor_scaling_of(
                    [&](int gas) { return in.col(gas, cd); },
                    md.gas, md.dens != 0, md.scal_gas, md.compl_ != 0, k.idx_h2o, pv, tv, cd);
        }
        st.scaling[q] = sc;
    }
}

template<typename Real, bool LW, typename In>
__device__ __forceinline__ void ps_aux(const Kd<Real>& k, int ibnd, const In& in, Real tv, Real cd, bool at_sfc,
                                       PairState<Real>& st)
{
    if constexpr (LW)
    {
        const Real* tab = k.totplnk + int64_t(ibnd) * k.nplanck;
        st.aux[0] = interpolate1D(tv, k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        st.aux[1] = interpolate1D(in.t_hi(), k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        st.aux[2] = interpolate1D(in.t_lo(), k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        if (at_sfc)
        {
            const Real ts = in.t_sfc();
            st.aux[3] = interpolate1D(ts, k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
            st.aux[4] = interpolate1D(ts + Real(1), k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        }
    }
    else
        st.aux[0] = in.col(k.idx_h2o, cd) + cd;
}

template<typename Real, bool LW, typename In>
__device__ __forceinline__ void pair_state_in(const Kd<Real>& k, const GoBand& bd, int ibnd, const In& in,
                                              bool at_sfc, PairState<Real>& st)
{
    const Real pv = in.p(), tv = in.t(), cd = in.cd();
    const CL<Real> c = interp_cl(k, pv, tv);
    ps_major(k, bd, c, in, cd, st);
    ps_minor(k, bd, c.itropo, in, pv, tv, cd, 0, 1, st);
    ps_aux<Real, LW>(k, ibnd, in, tv, cd, at_sfc, st);
}

// Band-independent state of a column, computed once per work-group by the fused kernels and kept
// in LDS (element [i][column]): ci = jtemp, jpress, itropo (+ LW: Planck-table index of T_lay, T_hi,
// T_lo); cf = ftemp, fpress (+ LW: the Planck-table fractions).
template<typename Real, bool LW>
struct ColState
{
    int ci[LW ? 6 : 3][GO_COLS];
    Real cf[LW ? 5 : 2][GO_COLS];
};

// Fills ColState for the 64 columns (wavefront 0: the interpolation; LW, wavefront 1: the Planck
// positions); the caller synchronises before it is read.
template<typename Real, bool LW, typename In>
__device__ __forceinline__ void col_state(const Kd<Real>& k, const In& in, int w, int lane, ColState<Real, LW>& cs)
{
    if (w == 0)
    {
        const CL<Real> c = interp_cl(k, in.p(), in.t());
        cs.ci[0][lane] = c.jtemp; cs.ci[1][lane] = c.jpress; cs.ci[2][lane] = c.itropo;
        cs.cf[0][lane] = c.ftemp; cs.cf[1][lane] = c.fpress;
    }
    else if constexpr (LW)
    {
        if (w == 1)
        {
            const Real t[3] = {in.t(), in.t_hi(), in.t_lo()};
#pragma unroll
            for (int q = 0; q < 3; ++q)
                interp1d_pos(t[q], k.temp_ref_min, k.totplnk_delta, k.nplanck, cs.ci[3 + q][lane], cs.cf[2 + q][lane]);
        }
    }
}

// Phase 1 of the fused forward kernels on wavefront w (uniform) of four, every wavefront taking all
// 64 columns: w = 0 the major interpolation state; LW: w = 2 the Planck values, w = 1 and 3 the even
// and odd minor slots; SW: w = 1, 2, 3 every third minor slot, w = 1 also the Rayleigh column.  The
// parts write disjoint fields of the pair state.
template<typename Real, bool LW, typename In>
__device__ __forceinline__ void pair_state_split(const Kd<Real>& k, const GoBand& bd, int ibnd, const In& in,
                                                 const ColState<Real, LW>& cs, int lane, bool at_sfc,
                                                 PairState<Real>& st, int w)
{
    const Real pv = in.p(), tv = in.t(), cd = in.cd();
    const CL<Real> c{cs.ci[0][lane], cs.ci[1][lane], cs.ci[2][lane], cs.cf[0][lane], cs.cf[1][lane]};
    if (w == 0)
        ps_major(k, bd, c, in, cd, st);
    else if constexpr (LW)
    {
        if (w == 2)
        {
            const Real* tab = k.totplnk + int64_t(ibnd) * k.nplanck;
#pragma unroll
            for (int q = 0; q < 3; ++q) st.aux[q] = interp1d_at(tab, cs.ci[3 + q][lane], cs.cf[2 + q][lane]);
            if (at_sfc)
            {
                const Real ts = in.t_sfc();
                st.aux[3] = interpolate1D(ts, k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
                st.aux[4] = interpolate1D(ts + Real(1), k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
            }
        }
        else
            ps_minor(k, bd, c.itropo, in, pv, tv, cd, w == 1 ? 0 : 1, 2, st);
    }
    else
    {
        ps_minor(k, bd, c.itropo, in, pv, tv, cd, w - 1, 3, st);
        if (w == 1) ps_aux<Real, false>(k, ibnd, in, tv, cd, false, st);
    }
}

template<typename Real, bool LW>
__device__ __forceinline__ void pair_state(const Kd<Real>& k, const GasRef<Real>* sg, const GoBand& bd, int ibnd,
                                           int64_t icol, int64_t ilay, int64_t ncol, const Real* play,
                                           const Real* tlay, const Real* tlev, const Real* tsfc,
                                           const Real* col_dry_a, bool at_sfc, PairState<Real>& st)
{
    const PairInGlobal<Real> in{sg, icol, ilay, ncol, play, tlay, tlev, tsfc, col_dry_a};
    pair_state_in<Real, LW>(k, bd, ibnd, in, at_sfc, st);
}

// Entries i..i+3 of a band-packed table (quads are 4-element aligned).  The table base is
// uniform and the byte offset fits 32 bits (kdist_pack checks it), so the read is addressed as
// scalar base + 32-bit lane offset.
template<typename Real>
__device__ __forceinline__ void load4(const Real* __restrict__ base, uint32_t i, Real (&v)[4])
{
    const char* p = reinterpret_cast<const char*>(base) + (i * uint32_t(sizeof(Real)));
    if constexpr (sizeof(Real) == 4)
    {
        const float4 a = *reinterpret_cast<const float4*>(p);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
    else
    {
        const double2 a = *reinterpret_cast<const double2*>(p);
        const double2 b = *reinterpret_cast<const double2*>(p + 16);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
}

// Phase 2: major + minor absorption of g-points gl..gl+3 of the band (same arithmetic order
// per g-point as major_tau and the minor sum of tau_absorption_kernel).
template<typename Real>
__device__ __forceinline__ void absorption4(const Kd<Real>& k, const GoBand& bd, const PairState<Real>& st,
                                            uint32_t gl, Real (&tau)[4])
{
    const uint32_t gp = bd.gp;
    // Minor sum first: its loop then holds no major-gather data live.
    Real minor[4] = {Real(0), Real(0), Real(0), Real(0)};
    const int nslot = bd.nslot;
    const uint32_t sQ = uint32_t(nslot) * gp;          // eta stride
    const uint32_t mb = uint32_t(bd.moff) + gl;
    for (int q = 0; q < nslot; ++q)
    {
        Real a[4], b[4], c[4], d[4];
        const uint32_t o0 = mb + uint32_t(st.mn[0]) + uint32_t(q) * gp;
        const uint32_t o1 = mb + uint32_t(st.mn[1]) + uint32_t(q) * gp;
        load4(k.kminor_p, o0, a); load4(k.kminor_p, o0 + sQ, b);
        load4(k.kminor_p, o1, c); load4(k.kminor_p, o1 + sQ, d);
        const Real sc = st.scaling[q];
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const Real lt = st.fminor[0][0] * a[e] + st.fminor[0][1] * b[e] +
                            st.fminor[1][0] * c[e] + st.fminor[1][1] * d[e];
            minor[e] += lt * sc;
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t sE = uint32_t(k.npres + 1) * gp, sP = gp;
    const uint32_t kb = uint32_t(bd.koff) + gl;
#pragma unroll
    for (int e = 0; e < 4; ++e) tau[e] = Real(0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        Real a[4], b[4], c[4], d[4];
        const uint32_t o = kb + uint32_t(st.om[i]);
        load4(k.kmajor_p, o, a); load4(k.kmajor_p, o + sE, b);
        load4(k.kmajor_p, o + sP, c); load4(k.kmajor_p, o + sP + sE, d);
#pragma unroll
        for (int e = 0; e < 4; ++e)
            tau[e] += st.col_mix[i] * (st.fmajor[i][0] * a[e] + st.fmajor[i][1] * b[e] +
                                       st.fmajor[i][2] * c[e] + st.fmajor[i][3] * d[e]);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) tau[e] += minor[e];
}

// Planck fraction of g-points gl..gl+3 (.cu:170-180): the corners and weights of the major gather.
template<typename Real>
__device__ __forceinline__ void planck_fraction4(const Kd<Real>& k, const GoBand& bd, const PairState<Real>& st,
                                                 uint32_t gl, Real (&pf)[4])
{
    const uint32_t gp = bd.gp, sE = uint32_t(k.npres + 1) * gp, sP = gp;
    const uint32_t kb = uint32_t(bd.koff) + gl;
    Real v[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        Real a[4], b[4], c[4], d[4];
        const uint32_t o = kb + uint32_t(st.om[i]);
        load4(k.planck_frac_p, o, a); load4(k.planck_frac_p, o + sE, b);
        load4(k.planck_frac_p, o + sP, c); load4(k.planck_frac_p, o + sP + sE, d);
#pragma unroll
        for (int e = 0; e < 4; ++e)
            v[i][e] = st.fmajor[i][0] * a[e] + st.fmajor[i][1] * b[e] + st.fmajor[i][2] * c[e] +
                      st.fmajor[i][3] * d[e];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) pf[e] = v[0][e] + v[1][e];
}

// Rayleigh optical depth of g-points gl..gl+3 (.cu:572-616).
template<typename Real>
__device__ __forceinline__ void rayleigh4(const Kd<Real>& k, const GoBand& bd, const PairState<Real>& st,
                                          uint32_t gl, Real (&tr)[4])
{
    const uint32_t gp = bd.gp;
    const uint32_t rb = uint32_t(bd.roff) + gl;
    Real a[4], b[4], c[4], d[4];
    load4(k.krayl_p, rb + uint32_t(st.ry[0]), a); load4(k.krayl_p, rb + uint32_t(st.ry[0]) + gp, b);
    load4(k.krayl_p, rb + uint32_t(st.ry[1]), c); load4(k.krayl_p, rb + uint32_t(st.ry[1]) + gp, d);
#pragma unroll
    for (int e = 0; e < 4; ++e)
    {
        const Real kloc = st.fminor[0][0] * a[e] + st.fminor[0][1] * b[e] +
                          st.fminor[1][0] * c[e] + st.fminor[1][1] * d[e];
        tr[e] = kloc * st.aux[0];
    }
}

// ---------------------------------------------------------------------------------------------
// Fused longwave gas optics: tau (+ cloud absorption per band) and Planck sources.  A work-group
// takes one (64-column tile, layer) unit: it stages the columns' inputs in LDS once, then runs every
// 16-g-point chunk of the launch's bands in three phases:
//   phase 1 (first chunk of a band; lane = column, the parts of the pair state spread over the four
//   wavefronts): pair states -> LDS;
//   phase 2 (lane = (pair, quad of g-points)): absorption and Planck fraction -> LDS;
//   phase 3 (lane = column, wavefront = g-point row): sources, cloud add_to, coalesced stores.
template<typename Real>
__global__ __launch_bounds__(GO_THREADS) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 7 : 4)))
void gas_optics_lw_kernel(Kd<Real> k, Gases<Real> g, int64_t ncol, int64_t nlay, GoGrid grid,
                          const Real* __restrict__ play, const Real* __restrict__ tlay,
                          const Real* __restrict__ tlev, const Real* __restrict__ tsfc,
                          const Real* __restrict__ col_dry_a, int sfc_lay,
                          const Real* __restrict__ cloud_tau, Real* __restrict__ tau,
                          Real* __restrict__ lay_src, Real* __restrict__ lev_inc,
                          Real* __restrict__ lev_dec, Real* __restrict__ sfc_src,
                          Real* __restrict__ sfc_jac, Real* __restrict__ bnd_lay,
                          Real* __restrict__ bnd_lev)
{
    static_assert(GO_GPTS == 16 && GO_THREADS == 4 * GO_COLS, "lane layout of phase 2");
    __shared__ PairState<Real> sst[GO_COLS];
    __shared__ __align__(16) Real s_tau[GO_GPTS][GO_COLS];
    __shared__ Real s_pf[GO_GPTS][GO_COLS];
    __shared__ ColState<Real, true> s_cs;
    // The gas pointers are read by the input staging only: they share s_tau's LDS.
    static_assert(sizeof(GasRef<Real>) * RRTMGP_MAX_GAS <= sizeof(Real) * GO_GPTS * GO_COLS, "sg in s_tau");
    GasRef<Real>* const sg = reinterpret_cast<GasRef<Real>*>(&s_tau[0][0]);
    extern __shared__ __align__(16) unsigned char go_dyn[];
    Real* const s_in = reinterpret_cast<Real*>(go_dyn);      // phase 1's input rows (launch-sized)
    const int64_t u = xcd_remap(blockIdx.x, grid.ntl);
    if (u >= grid.ntl) return;                          // uniform over the work-group
    const GoTl tl = go_tl(grid, u);
    if (!tl.ok) return;                                 // uniform: a listed id outside the grid
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
    const int64_t ilay = tl.ilay, ncl = ncol * nlay;
    const bool at_sfc = (ilay == sfc_lay - 1);

    if (threadIdx.x == 0) stage_gases(g, k.ngas, sg);
    __syncthreads();
    stage_pair_inputs<Real, true>(sg, k.ngas, tl.col0, ilay, ncol, nlay, play, tlay, tlev, tsfc, col_dry_a, s_in);
    __syncthreads();
    col_state<Real, true>(k, PairInLds<Real>{s_in, lane, k.ngas}, w, lane, s_cs);
    __syncthreads();
    // Every chunk (16 g-points of a band) of the launch's bands for this tile and layer: phase 1's
    // inputs are loaded once and the work-group's launch and staging are shared by all of them.
    const int nchunk = grid.nbnd * int(grid.cpb);
    for (int cb = 0; cb < nchunk; ++cb)
    {
        const int ln = lane;
        const int64_t icol = tl.col0 + ln;
        const int64_t idx = icol + ncol * ilay;
        const int ibnd = grid.b0 + cb / int(grid.cpb);
        const GoBand& bd = k.bands[ibnd];
        const int c0 = GO_GPTS * (cb % int(grid.cpb));
        const int ncg = min(bd.ng - c0, GO_GPTS);
        if (ncg <= 0) continue;                         // uniform: an empty chunk of a short band
        if (c0 == 0)                                    // pair states hold for all of a band's chunks
        {
            if (icol >= ncol)
            {
                if (w == 0) sst[ln].flags = 0;
            }
            else
                pair_state_split<Real, true>(k, bd, ibnd, PairInLds<Real>{s_in, ln, k.ngas}, s_cs, ln, at_sfc,
                                             sst[ln], w);
        }
        __syncthreads();

        // The band cloud optical depth of phase 3's column, requested before the gathers.
        const Real ctau = (cloud_tau && icol < ncol) ? cloud_tau[idx + ncl * ibnd] : Real(0);
        {
            const int pi = w * 16 + ln / 4;           // pair
            const int gq = 4 * (ln % 4);              // first g-point of the quad in the chunk
            const PairState<Real>& st = sst[pi];
            Real t4[4] = {Real(0), Real(0), Real(0), Real(0)};
            Real p4[4] = {Real(0), Real(0), Real(0), Real(0)};
            if ((st.flags & 1) && gq < ncg)
            {
                const uint32_t gl = uint32_t(c0 + gq);
                absorption4(k, bd, st, gl, t4);
                // Bounds the table reads in flight (and so the registers) to one gather group.
                __builtin_amdgcn_sched_barrier(0);
                planck_fraction4(k, bd, st, gl, p4);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                s_tau[gq + e][go_swz(gq + e, pi)] = t4[e];
                s_pf[gq + e][go_swz(gq + e, pi)] = p4[e];
            }
        }
        __syncthreads();

        if (icol < ncol)
        {
            const PairState<Real>& st = sst[ln];
            const int gbase = bd.g0 - k.bands[grid.b0].g0 + c0;
            // The g-point outputs are streamed (non-temporal stores): the solvers read them back
            // long after any cache could hold them.
            if (bnd_lay)
            {
                // Factored sources: Planck fractions per g-point, band Planck values per layer and
                // level (written once, by the band's first chunk; level nlay by the top layer).
                if (w == 0 && c0 == 0)
                {
                    const int64_t nlev = nlay + 1, bb = ibnd - grid.b0;
                    bnd_lay[idx + ncl * bb] = st.aux[0];
                    bnd_lev[icol + ncol * ilay + ncol * nlev * bb] = st.aux[2];
                    if (ilay == nlay - 1) bnd_lev[icol + ncol * nlay + ncol * nlev * bb] = st.aux[1];
                }
                for (int row = w; row < ncg; row += GO_THREADS / 64)
                {
                    const int gr = gbase + row;
                    const int64_t o = idx + ncl * gr;
                    const Real pf = s_pf[row][go_swz(row, ln)];
                    const Real tv = s_tau[row][go_swz(row, ln)];
                    __builtin_nontemporal_store(cloud_tau ? tv + ctau : tv, &tau[o]);
                    __builtin_nontemporal_store(pf, &lay_src[o]);
                    if (at_sfc)
                    {
                        sfc_src[icol + ncol * gr] = pf * st.aux[3];
                        sfc_jac[icol + ncol * gr] = pf * (st.aux[4] - st.aux[3]);
                    }
                }
            }
            else
                for (int row = w; row < ncg; row += GO_THREADS / 64)
                {
                    const int gr = gbase + row;
                    const int64_t o = idx + ncl * gr;
                    const Real pf = s_pf[row][go_swz(row, ln)];
                    const Real tv = s_tau[row][go_swz(row, ln)];
                    __builtin_nontemporal_store(cloud_tau ? tv + ctau : tv, &tau[o]);
                    __builtin_nontemporal_store(pf * st.aux[0], &lay_src[o]);
                    __builtin_nontemporal_store(pf * st.aux[1], &lev_inc[o]);
                    __builtin_nontemporal_store(pf * st.aux[2], &lev_dec[o]);
                    if (at_sfc)
                    {
                        sfc_src[icol + ncol * gr] = pf * st.aux[3];
                        sfc_jac[icol + ncol * gr] = pf * (st.aux[4] - st.aux[3]);
                    }
                }
        }
        __syncthreads();                                // the next chunk rewrites sst and the rows
    }
}

// Fused shortwave gas optics: absorption + Rayleigh + combine (+ band cloud add_to), in the
// same three phases.
template<typename Real>
__global__ __launch_bounds__(GO_THREADS) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 8 : 4)))
void gas_optics_sw_kernel(Kd<Real> k, Gases<Real> g, int64_t ncol, int64_t nlay, GoGrid grid,
                          const Real* __restrict__ play, const Real* __restrict__ tlay,
                          const Real* __restrict__ col_dry_a, const Real* __restrict__ ctau_a,
                          const Real* __restrict__ cssa_a, const Real* __restrict__ cg_a,
                          Real* __restrict__ tau, Real* __restrict__ ssa, Real* __restrict__ gout, int coalb)
{
    static_assert(GO_GPTS == 16 && GO_THREADS == 4 * GO_COLS, "lane layout of phase 2");
    __shared__ PairState<Real> sst[GO_COLS];
    __shared__ __align__(16) Real s_abs[GO_GPTS][GO_COLS];
    __shared__ Real s_ray[GO_GPTS][GO_COLS];
    __shared__ ColState<Real, false> s_cs;
    // The gas pointers are read by the input staging only: they share s_abs's LDS.
    static_assert(sizeof(GasRef<Real>) * RRTMGP_MAX_GAS <= sizeof(Real) * GO_GPTS * GO_COLS, "sg in s_abs");
    GasRef<Real>* const sg = reinterpret_cast<GasRef<Real>*>(&s_abs[0][0]);
    extern __shared__ __align__(16) unsigned char go_dyn[];
    Real* const s_in = reinterpret_cast<Real*>(go_dyn);      // phase 1's input rows (launch-sized)
    const int64_t u = xcd_remap(blockIdx.x, grid.ntl);
    if (u >= grid.ntl) return;
    const GoTl tl = go_tl(grid, u);
    if (!tl.ok) return;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
    const int64_t ilay = tl.ilay, ncl = ncol * nlay;
    const Real eps2 = Real(2 * prop_eps);
    const Real op_eps = Real(3) * Limits<Real>::tiny;

    if (threadIdx.x == 0) stage_gases(g, k.ngas, sg);
    __syncthreads();
    stage_pair_inputs<Real, false>(sg, k.ngas, tl.col0, ilay, ncol, nlay, play, tlay, nullptr, nullptr, col_dry_a,
                                   s_in);
    __syncthreads();
    col_state<Real, false>(k, PairInLds<Real>{s_in, lane, k.ngas}, w, lane, s_cs);
    __syncthreads();
    const int nchunk = grid.nbnd * int(grid.cpb);
    for (int cb = 0; cb < nchunk; ++cb)
    {
        const int ln = lane;
        const int64_t icol = tl.col0 + ln;
        const int64_t idx = icol + ncol * ilay;
        const int ibnd = grid.b0 + cb / int(grid.cpb);
        const GoBand& bd = k.bands[ibnd];
        const int c0 = GO_GPTS * (cb % int(grid.cpb));
        const int ncg = min(bd.ng - c0, GO_GPTS);
        if (ncg <= 0) continue;
        if (c0 == 0)                                    // pair states hold for all of a band's chunks
        {
            if (icol >= ncol)
            {
                if (w == 0) sst[ln].flags = 0;
            }
            else
                pair_state_split<Real, false>(k, bd, ibnd, PairInLds<Real>{s_in, ln, k.ngas}, s_cs, ln, false,
                                              sst[ln], w);
        }
        __syncthreads();

        // The band cloud properties of phase 3's column, requested before the gathers.
        Real t2 = Real(0), s2 = Real(0), g2 = Real(0);
        if (ctau_a && icol < ncol)
        {
            const int64_t ib = idx + ncl * ibnd;
            t2 = ctau_a[ib]; s2 = cssa_a[ib]; g2 = cg_a[ib];
        }
        {
            const int pi = w * 16 + ln / 4;
            const int gq = 4 * (ln % 4);
            const PairState<Real>& st = sst[pi];
            Real t4[4] = {Real(0), Real(0), Real(0), Real(0)};
            Real r4[4] = {Real(0), Real(0), Real(0), Real(0)};
            if ((st.flags & 1) && gq < ncg)
            {
                const uint32_t gl = uint32_t(c0 + gq);
                absorption4(k, bd, st, gl, t4);
                __builtin_amdgcn_sched_barrier(0);
                rayleigh4(k, bd, st, gl, r4);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                s_abs[gq + e][go_swz(gq + e, pi)] = t4[e];
                s_ray[gq + e][go_swz(gq + e, pi)] = r4[e];
            }
        }
        __syncthreads();

        if (icol < ncol)
        {
            const int gbase = bd.g0 - k.bands[grid.b0].g0 + c0;
            for (int row = w; row < ncg; row += GO_THREADS / 64)
            {
                const Real tabs = s_abs[row][go_swz(row, ln)];
                const Real tray = s_ray[row][go_swz(row, ln)];
                const Real tt = tabs + tray;
                const int64_t o = idx + ncl * (gbase + row);
                if (coalb)
                {
                    // The same with the co-albedo 1 - ssa in place of ssa, for the cloud input and
                    // the output: formed from the absorption optical depths, it keeps the precision
                    // of the absorption that 1 - ssa in FP32 loses once it falls below 6e-8.
                    Real tt_out = tt;
                    Real co = (tt > eps2) ? tabs / tt : Real(1);
                    Real gg_out = Real(0);
                    if (ctau_a)
                    {
                        const Real tau12 = tt + t2;
                        const Real tauscat12 = (tt * (Real(1) - co)) + (t2 * (Real(1) - s2));
                        gg_out = (t2 * (Real(1) - s2) * g2) / max(tauscat12, op_eps);
                        co = (tau12 > op_eps) ? ((tt * co) + (t2 * s2)) / tau12 : Real(1) - tauscat12 / op_eps;
                        tt_out = tau12;
                    }
                    __builtin_nontemporal_store(tt_out, &tau[o]);
                    __builtin_nontemporal_store(co, &ssa[o]);
                    __builtin_nontemporal_store(gg_out, &gout[o]);
                    continue;
                }
                // The CPU combine rule (Gas_optics_rrtmgp.cpp:365-385).
                Real tt_out = tt;
                Real ss = (tt > eps2) ? tray / tt : Real(0);
                Real gg_out = Real(0);
                if (ctau_a)
                {
                    // inc_2stream_by_2stream_bybnd (optical_props_kernels.cu:99-136).
                    const Real tau12 = tt + t2;
                    const Real tauscat12 = (tt * ss) + (t2 * s2);
                    gg_out = ((tt * ss * Real(0)) + (t2 * s2 * g2)) / max(tauscat12, op_eps);
                    ss = tauscat12 / max(op_eps, tau12);
                    tt_out = tau12;
                }
                __builtin_nontemporal_store(tt_out, &tau[o]);
                __builtin_nontemporal_store(ss, &ssa[o]);
                __builtin_nontemporal_store(gg_out, &gout[o]);
            }
        }
        __syncthreads();                                // the next chunk rewrites sst and the rows
    }
}

// ---------------------------------------------------------------------------------------------
// Gas-optics adjoint.  For one band, given the cotangents of the band's fused outputs, adds the
// gradients with respect to the per-(column, layer) inputs: the column amounts col_gas
// (gas 0 = dry air), T_lay, ln p_lay and, in the longwave, T_lev at the layer's two levels and
// T_sfc.  Per g-point the outputs are linear in the table values with the pair's weights:
//   tau_abs = sum_i col_mix_i sum_c fmajor_ic K_icg + sum_q scaling_q sum_ie fminor_ie Km_qieg
//   pfrac   = sum_ic fmajor_ic P_icg,    tau_ray = col_ray sum_ie fminor_ie Kr_ieg,
// so phase 2 accumulates, per pair, the gradients of those weights over the band's g-points
// (one quad of g-points per lane, reduced over the pair's four lanes); phase 3 carries them
// through the interpolation (ftemp, fpress, eta -> fmajor/fminor, col_mix), the minor scalings and
// the Planck interpolation by hand.  The integer indices are piecewise constant.

constexpr int GA_CM = 0, GA_FM = 2, GA_SC = 10, GA_FN = 18, GA_X = 22, GA_N = 26;

// Derivative of interpolate1D with respect to val.
template<typename Real>
__device__ __forceinline__ Real interpolate1D_slope(Real val, Real offset, Real delta, int len, const Real* table)
{
    const Real val0 = (val - offset) / delta;
    const int idx = min(len - 1, max(1, int(val0) + 1));
    return (table[idx] - table[idx - 1]) / delta;
}

// Phase 3: the gradients of one pair's weights (acc) carried to the inputs.
template<typename Real, bool LW>
__device__ __forceinline__ void pair_state_backward(const Kd<Real>& k, const GasRef<Real>* sg, const GoBand& bd,
                                                    int ibnd, int64_t icol, int64_t ilay, int64_t ncol, int64_t ncl,
                                                    const Real* play, const Real* tlay, const Real* tlev,
                                                    const Real* tsfc, const Real* col_dry_a, bool at_sfc,
                                                    const Real* acc, Real* d_col, Real* d_tlay, Real* d_lnp,
                                                    Real* d_tlev_lo, Real* d_tlev_hi, Real* d_tsfc)
{
    const int64_t idx = icol + ncol * ilay;
    const Real pv = play[idx], tv = tlay[idx], cd = col_dry_a[idx];
    const CL<Real> c = interp_cl(k, pv, tv);
    const int s = c.itropo;
    const int gas1 = bd.gas[s][0], gas2 = bd.gas[s][1];
    const Real c1 = col_lds(sg, gas1, icol, ilay, cd), c2 = col_lds(sg, gas2, icol, ilay, cd);
    auto dcol = [&](int gas, Real v) { d_col[idx + ncl * gas] += v; };
    // g_dry: the derivative with respect to col_dry at fixed volume mixing ratios, times col_dry.
    // Every column the gas optics reads is a vmr times col_dry, and each quantity formed from
    // them is homogeneous in col_dry: col_mix, the minor scalings and col_ray of degree 1, eta and
    // the scaling fraction X of degree 0.  So g_dry = sum over the degree-1 quantities of their
    // cotangent times their value; the eta and X paths, whose per-gas terms cancel exactly in
    // the sum over gases but not in rounded arithmetic, do not enter.
    Real g_dry = Real(0);
    Real g_ftemp = Real(0), g_fp = Real(0), g_c1 = Real(0), g_c2 = Real(0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        // interp_flav (.cu:254-291) for itemp = i, with its intermediates.
        const int64_t vb = c.itropo + int64_t(c.jtemp + i - 1) * (k.ngas + 1) * 2;
        const Real r = k.vmr_ref[vb + 2 * gas1] / k.vmr_ref[vb + 2 * gas2];
        const Real cm = c1 + r * c2;
        const bool mix = cm > Real(2) * Limits<Real>::tiny;
        const Real eta = mix ? c1 / cm : Real(0.5);
        const Real feta = frac1(eta * Real(k.neta - 1));
        const Real ftt = i == 0 ? Real(1) - c.ftemp : c.ftemp;
        const Real m0 = (Real(1) - feta) * ftt, m1 = feta * ftt;
        const Real* gfm = acc + GA_FM + 4 * i;
        const Real* gfn = acc + GA_FN + 2 * i;
        const Real g_m0 = (Real(1) - c.fpress) * gfm[0] + c.fpress * gfm[2] + gfn[0];
        const Real g_m1 = (Real(1) - c.fpress) * gfm[1] + c.fpress * gfm[3] + gfn[1];
        g_fp += -m0 * gfm[0] - m1 * gfm[1] + m0 * gfm[2] + m1 * gfm[3];
        const Real g_feta = ftt * (g_m1 - g_m0);
        const Real g_ftt = (Real(1) - feta) * g_m0 + feta * g_m1;
        g_ftemp += i == 0 ? -g_ftt : g_ftt;
        const Real g_cm = acc[GA_CM + i];
        g_c1 += g_cm;
        g_c2 += g_cm * r;
        g_dry += g_cm * cm;
        if (mix)
        {
            // d eta / d c1 = r c2 / cm^2 = (1 - eta) / cm and d eta / d c2 = -eta r / cm, formed without
            // squaring cm: column amounts reach ~1e24 molecules cm-2, whose square overflows FP32.
            const Real g_eta = g_feta * Real(k.neta - 1);
            const Real icm = Real(1) / cm;
            g_c1 += g_eta * (r * c2 * icm) * icm;
            g_c2 -= g_eta * eta * r * icm;
        }
    }
    dcol(gas1, g_c1);
    dcol(gas2, g_c2);
    Real g_t = g_ftemp / k.temp_ref_delta;
    Real g_lnp = g_fp / k.press_ref_log_delta;
    // Minor scalings (.cu:388-420): sc = col(gas) [P F], P = 0.01 p / T, F = X or 1 - X with
    // X = col(scal) / (col_dry + col(h2o)).
    const int nq = bd.nq[s];
    for (int q = 0; q < nq; ++q)
    {
        const GoMinor md = bd.minor[s][q];
        const Real g_sc = acc[GA_SC + q];
        const Real base = col_lds(sg, md.gas, icol, ilay, cd);
        if (!md.dens)
        {
            dcol(md.gas, g_sc);
            g_dry += g_sc * base;
            continue;
        }
        const Real P = Real(0.01) * pv / tv;
        Real F = Real(1), X = Real(0), den = Real(1);
        if (md.scal_gas > 0)
        {
            den = cd + col_lds(sg, k.idx_h2o, icol, ilay, cd);
            X = col_lds(sg, md.scal_gas, icol, ilay, cd) / den;
            F = md.compl_ ? Real(1) - X : X;
        }
        dcol(md.gas, g_sc * P * F);
        g_dry += g_sc * base * P * F;
        const Real g_P = g_sc * base * F;
        g_lnp += g_P * P;
        g_t -= g_P * P / tv;
        if (md.scal_gas > 0)
        {
            const Real g_X = (md.compl_ ? -g_sc : g_sc) * base * P;
            dcol(md.scal_gas, g_X / den);
            dcol(0, -g_X * X / den);
            dcol(k.idx_h2o, -g_X * X / den);
        }
    }
    if constexpr (LW)
    {
        const Real* tab = k.totplnk + int64_t(ibnd) * k.nplanck;
        g_t += acc[GA_X + 0] * interpolate1D_slope(tv, k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        d_tlev_hi[idx] += acc[GA_X + 1] * interpolate1D_slope(tlev[icol + ncol * (ilay + 1)], k.temp_ref_min,
                                                              k.totplnk_delta, k.nplanck, tab);
        d_tlev_lo[idx] += acc[GA_X + 2] * interpolate1D_slope(tlev[idx], k.temp_ref_min, k.totplnk_delta,
                                                              k.nplanck, tab);
        if (at_sfc)
            d_tsfc[icol] += acc[GA_X + 3] * interpolate1D_slope(tsfc[icol], k.temp_ref_min, k.totplnk_delta,
                                                                k.nplanck, tab);
    }
    else
    {
        dcol(k.idx_h2o, acc[GA_X]);      // col_ray = col(h2o) + col_dry
        dcol(0, acc[GA_X]);
        g_dry += acc[GA_X] * (col_lds(sg, k.idx_h2o, icol, ilay, cd) + cd);
    }
    dcol(k.ngas + 1, g_dry / cd);
    d_tlay[idx] += g_t;
    d_lnp[idx] += g_lnp;
}

// Sum over the four lanes of a pair (lanes 4p .. 4p+3 of a wavefront).
__device__ __forceinline__ float quad_sum(float v)
{
    v += __shfl_xor(v, 1);
    return v + __shfl_xor(v, 2);
}
__device__ __forceinline__ double quad_sum(double v)
{
    v += __shfl_xor(v, 1);
    return v + __shfl_xor(v, 2);
}

// One band; a work-group takes 64 columns of one layer and loops over the band's 16-g-point
// chunks.  LW cotangents: ga (tau), gb (lay_src), gc (lev_src_inc), gd (lev_src_dec) of
// (ncol, nlay, ngb) and gs (sfc_src, (ncol, ngb)); SW: ga (tau), gb (ssa), gc (g).
template<typename Real, bool LW>
__global__ __launch_bounds__(GO_THREADS) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 3 : 1)))
void gas_optics_adjoint_kernel(Kd<Real> k, Gases<Real> g, int64_t ncol, int64_t nlay, int ibnd,
                               const Real* __restrict__ play, const Real* __restrict__ tlay,
                               const Real* __restrict__ tlev, const Real* __restrict__ tsfc,
                               const Real* __restrict__ col_dry_a, int sfc_lay,
                               const Real* __restrict__ ctau_a, const Real* __restrict__ cssa_a,
                               const Real* __restrict__ cg_a,
                               const Real* __restrict__ ga, const Real* __restrict__ gb,
                               const Real* __restrict__ gc, const Real* __restrict__ gd,
                               const Real* __restrict__ gs,
                               Real* __restrict__ d_col, Real* __restrict__ d_tlay, Real* __restrict__ d_lnp,
                               Real* __restrict__ d_tlev_lo, Real* __restrict__ d_tlev_hi, Real* __restrict__ d_tsfc)
{
    static_assert(GO_GPTS == 16 && GO_THREADS == 4 * GO_COLS, "lane layout of phase 2");
    __shared__ PairState<Real> sst[GO_COLS];
    __shared__ Real s_g[5][GO_GPTS][GO_ROW];
    __shared__ Real s_acc[GO_COLS][GA_N + 1];
    __shared__ Real s_cl[3][GO_COLS];
    __shared__ Real s_sc[GO_MAXM][GO_THREADS];      // per lane: cotangent sums of the minor scalings
    __shared__ GasRef<Real> sg[RRTMGP_MAX_GAS];
    const int64_t ncoltiles = (ncol + GO_COLS - 1) / GO_COLS;
    const int64_t u = xcd_remap(blockIdx.x, ncoltiles * nlay);
    if (u >= ncoltiles * nlay) return;                     // uniform over the work-group
    const int64_t ilay = u / ncoltiles, col0 = (u % ncoltiles) * GO_COLS;
    const GoBand& bd = k.bands[ibnd];
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int64_t ncl = ncol * nlay;
    const bool at_sfc = LW && (ilay == sfc_lay - 1);

    if (threadIdx.x == 0) stage_gases(g, k.ngas, sg);
    __syncthreads();
    if (w == 0)
    {
        const int64_t icol = col0 + lane;
        if (icol < ncol)
        {
            pair_state<Real, LW>(k, sg, bd, ibnd, icol, ilay, ncol, play, tlay, tlev, tsfc, col_dry_a, at_sfc,
                                 sst[lane]);
            if (!LW)
            {
                const int64_t ib = icol + ncol * ilay + ncl * ibnd;
                s_cl[0][lane] = ctau_a ? ctau_a[ib] : Real(0);
                s_cl[1][lane] = ctau_a ? cssa_a[ib] : Real(0);
                s_cl[2][lane] = ctau_a ? cg_a[ib] : Real(0);
            }
        }
        else
            sst[lane].flags = 0;
    }

    const int pi = w * 16 + lane / 4;                  // pair of phase 2
    const int gq = 4 * (lane % 4);                     // first g-point of the lane's quad in a chunk
    Real acc[GA_N];
#pragma unroll
    for (int a = 0; a < GA_N; ++a) acc[a] = Real(0);
#pragma unroll
    for (int q = 0; q < GO_MAXM; ++q) s_sc[q][threadIdx.x] = Real(0);
    const Real eps2 = Real(2 * prop_eps);
    const Real op_eps = Real(3) * Limits<Real>::tiny;

    for (int c0 = 0; c0 < bd.ng; c0 += GO_GPTS)
    {
        const int ncg = min(bd.ng - c0, GO_GPTS);
        __syncthreads();                               // previous chunk's LDS reads are done
        {
            const int64_t icol = col0 + lane;
            for (int row = w; row < ncg; row += GO_THREADS / 64)
            {
                Real v0 = Real(0), v1 = Real(0), v2 = Real(0), v3 = Real(0), v4 = Real(0);
                if (icol < ncol)
                {
                    const int64_t o = icol + ncol * ilay + ncl * (c0 + row);
                    v0 = ga[o]; v1 = gb[o]; v2 = gc[o];
                    if (LW)
                    {
                        v3 = gd[o];
                        if (at_sfc) v4 = gs[icol + ncol * (c0 + row)];
                    }
                }
                s_g[0][row][lane] = v0; s_g[1][row][lane] = v1; s_g[2][row][lane] = v2;
                s_g[3][row][lane] = v3; s_g[4][row][lane] = v4;
            }
        }
        __syncthreads();
        const PairState<Real>& st = sst[pi];
        if (!((st.flags & 1) && gq < ncg)) continue;
        const uint32_t gl = uint32_t(c0 + gq);
        const uint32_t gp = bd.gp, sE = uint32_t(k.npres + 1) * gp, sP = gp;
        const uint32_t kb = uint32_t(bd.koff) + gl;
        Real K[2][4][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
        {
            const uint32_t o = kb + uint32_t(st.om[i]);
            load4(k.kmajor_p, o, K[i][0]); load4(k.kmajor_p, o + sE, K[i][1]);
            load4(k.kmajor_p, o + sP, K[i][2]); load4(k.kmajor_p, o + sP + sE, K[i][3]);
        }
        // Forward values per g-point that the cotangents need.
        Real gabs[4], gray[4];
        Real kr[2][2][4];
        if (!LW)
        {
            const uint32_t rb = uint32_t(bd.roff) + gl;
            load4(k.krayl_p, rb + uint32_t(st.ry[0]), kr[0][0]); load4(k.krayl_p, rb + uint32_t(st.ry[0]) + gp, kr[0][1]);
            load4(k.krayl_p, rb + uint32_t(st.ry[1]), kr[1][0]); load4(k.krayl_p, rb + uint32_t(st.ry[1]) + gp, kr[1][1]);
            Real tabs[4];
            absorption4(k, bd, st, gl, tabs);
            const Real t2 = s_cl[0][pi], s2 = s_cl[1][pi], g2 = s_cl[2][pi];
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                const Real tray = (st.fminor[0][0] * kr[0][0][e] + st.fminor[0][1] * kr[0][1][e] +
                                   st.fminor[1][0] * kr[1][0][e] + st.fminor[1][1] * kr[1][1][e]) * st.aux[0];
                const Real tt = tabs[e] + tray;
                const Real ss = (tt > eps2) ? tray / tt : Real(0);
                const Real tau12 = tt + t2;
                const Real tauscat = tt * ss + t2 * s2;
                const Real gout = (t2 * s2 * g2) / max(tauscat, op_eps);
                const Real ssa = tauscat / max(op_eps, tau12);
                const Real gto = s_g[0][gq + e][pi], gwo = s_g[1][gq + e][pi], ggo = s_g[2][gq + e][pi];
                const Real g_ts = (tau12 > op_eps ? gwo / tau12 : Real(0)) -
                                  (tauscat > op_eps ? ggo * gout / tauscat : Real(0));
                const Real g_t12 = gto - (tau12 > op_eps ? gwo * ssa / tau12 : Real(0));
                gabs[e] = g_t12;
                gray[e] = g_t12 + (tt > eps2 ? g_ts : Real(0));
            }
        }
        else
        {
#pragma unroll
            for (int e = 0; e < 4; ++e) { gabs[e] = s_g[0][gq + e][pi]; gray[e] = Real(0); }
        }
        // Major terms.
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
#pragma unroll
            for (int i = 0; i < 2; ++i)
            {
                Real sk = Real(0);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                {
                    sk += st.fmajor[i][c] * K[i][c][e];
                    acc[GA_FM + 4 * i + c] += gabs[e] * st.col_mix[i] * K[i][c][e];
                }
                acc[GA_CM + i] += gabs[e] * sk;
            }
            if (!LW)
            {
                Real sr = Real(0);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                    {
                        sr += st.fminor[i][c] * kr[i][c][e];
                        acc[GA_FN + 2 * i + c] += gray[e] * st.aux[0] * kr[i][c][e];
                    }
                acc[GA_X] += gray[e] * sr;
            }
        }
        // Planck-fraction terms (LW), after the major table values are dead: the two table slabs
        // are never live together.
        if constexpr (LW)
        {
            __builtin_amdgcn_sched_barrier(0);
            Real P[2][4][4];
#pragma unroll
            for (int i = 0; i < 2; ++i)
            {
                const uint32_t o = kb + uint32_t(st.om[i]);
                load4(k.planck_frac_p, o, P[i][0]); load4(k.planck_frac_p, o + sE, P[i][1]);
                load4(k.planck_frac_p, o + sP, P[i][2]); load4(k.planck_frac_p, o + sP + sE, P[i][3]);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                Real pf = Real(0);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int c = 0; c < 4; ++c) pf += st.fmajor[i][c] * P[i][c][e];
                const Real gl_ = s_g[1][gq + e][pi], gi_ = s_g[2][gq + e][pi], gd_ = s_g[3][gq + e][pi];
                const Real gs_ = at_sfc ? s_g[4][gq + e][pi] : Real(0);
                const Real gpf = gl_ * st.aux[0] + gi_ * st.aux[1] + gd_ * st.aux[2] +
                                 (at_sfc ? gs_ * st.aux[3] : Real(0));
                acc[GA_X + 0] += gl_ * pf; acc[GA_X + 1] += gi_ * pf;
                acc[GA_X + 2] += gd_ * pf; acc[GA_X + 3] += gs_ * pf;
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[GA_FM + 4 * i + c] += gpf * P[i][c][e];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // Minor terms, one slot at a time (a rolled loop keeps one slot's table values live); the
        // slot's scaling cotangent goes to the lane's LDS slot.
        const int nslot = bd.nslot;
        const uint32_t sQ = uint32_t(nslot) * gp;
        const uint32_t mb = uint32_t(bd.moff) + gl;
#pragma unroll 1
        for (int q = 0; q < nslot; ++q)
        {
            Real km[2][2][4];
            const uint32_t o0 = mb + uint32_t(st.mn[0]) + uint32_t(q) * gp;
            const uint32_t o1 = mb + uint32_t(st.mn[1]) + uint32_t(q) * gp;
            load4(k.kminor_p, o0, km[0][0]); load4(k.kminor_p, o0 + sQ, km[0][1]);
            load4(k.kminor_p, o1, km[1][0]); load4(k.kminor_p, o1 + sQ, km[1][1]);
            const Real scq = st.scaling[q];
            Real tq = Real(0);
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                Real sm = Real(0);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                    {
                        sm += st.fminor[i][c] * km[i][c][e];
                        acc[GA_FN + 2 * i + c] += gabs[e] * scq * km[i][c][e];
                    }
                tq += gabs[e] * sm;
            }
            s_sc[q][threadIdx.x] += tq;
        }
    }
    // Sum over the pair's four lanes; the first lane of each pair stores the pair's totals.
#pragma unroll
    for (int a = 0; a < GA_N; ++a)
    {
        const Real v = quad_sum(a >= GA_SC && a < GA_SC + GO_MAXM ? s_sc[a - GA_SC][threadIdx.x] : acc[a]);
        if (lane % 4 == 0) s_acc[pi][a] = v;
    }
    __syncthreads();
    if (w == 0)
    {
        const int64_t icol = col0 + lane;
        if (icol < ncol)
            pair_state_backward<Real, LW>(k, sg, bd, ibnd, icol, ilay, ncol, ncl, play, tlay, tlev, tsfc, col_dry_a,
                                          at_sfc, s_acc[lane], d_col, d_tlay, d_lnp, d_tlev_lo, d_tlev_hi, d_tsfc);
    }
}

// ---------------------------------------------------------------------------------------------
// API-parity kernels that materialize the interpolation state (reference launch structure).

template<typename Real>
__global__ void interpolation_kernel(Kd<Real> k, int64_t ncol, int64_t nlay,
                                     const Real* __restrict__ play, const Real* __restrict__ tlay,
                                     const Real* __restrict__ col_gas, int* __restrict__ jtemp,
                                     Real* __restrict__ fmajor, Real* __restrict__ fminor,
                                     Real* __restrict__ col_mix, Bool* __restrict__ tropo,
                                     int* __restrict__ jeta, int* __restrict__ jpress)
{
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ncl = ncol * nlay;
    if (idx >= ncl) return;
    const CL<Real> c = interp_cl(k, play[idx], tlay[idx]);
    jtemp[idx] = c.jtemp;
    jpress[idx] = c.jpress;
    tropo[idx] = Bool(c.itropo == 0);
    for (int iflav = 0; iflav < k.nflav; ++iflav)
    {
        const int gas1 = k.flavor[2 * iflav], gas2 = k.flavor[2 * iflav + 1];
        FL<Real> f;
        interp_flav(k, c, gas1, gas2, col_gas[idx + ncl * gas1], col_gas[idx + ncl * gas2], f);
        const int64_t fl = idx + ncl * iflav;
        for (int i = 0; i < 2; ++i)
        {
            jeta[i + 2 * fl] = f.jeta[i];
            col_mix[i + 2 * fl] = f.col_mix[i];
            fminor[2 * (i + 2 * fl)] = f.fminor[i][0];
            fminor[2 * (i + 2 * fl) + 1] = f.fminor[i][1];
            for (int q = 0; q < 4; ++q) fmajor[4 * (i + 2 * fl) + q] = f.fmajor[i][q];
        }
    }
}

template<typename Real>
__device__ __forceinline__ void load_state(const Kd<Real>& k, int64_t idx, int64_t ncl, int iflav,
                                           const Bool* tropo, const int* jtemp, const int* jpress,
                                           const int* jeta, const Real* col_mix, const Real* fminor,
                                           const Real* fmajor, CL<Real>& c, FL<Real>& f)
{
    c.itropo = !tropo[idx];
    c.jtemp = jtemp[idx];
    c.jpress = jpress ? jpress[idx] : 1;
    const int64_t fl = idx + ncl * iflav;
    for (int i = 0; i < 2; ++i)
    {
        f.jeta[i] = jeta[i + 2 * fl];
        f.col_mix[i] = col_mix ? col_mix[i + 2 * fl] : Real(0);
        f.fminor[i][0] = fminor ? fminor[2 * (i + 2 * fl)] : Real(0);
        f.fminor[i][1] = fminor ? fminor[2 * (i + 2 * fl) + 1] : Real(0);
        for (int q = 0; q < 4; ++q) f.fmajor[i][q] = fmajor ? fmajor[4 * (i + 2 * fl) + q] : Real(0);
    }
}

// compute_tau_absorption from materialized state: thread per (col, lay, band).
template<typename Real>
__global__ void tau_absorption_kernel(Kd<Real> k, int64_t ncol, int64_t nlay,
                                      const Bool* __restrict__ tropo, const Real* __restrict__ col_mix,
                                      const Real* __restrict__ fmajor, const Real* __restrict__ fminor,
                                      const Real* __restrict__ play, const Real* __restrict__ tlay,
                                      const Real* __restrict__ col_gas, const int* __restrict__ jeta,
                                      const int* __restrict__ jtemp, const int* __restrict__ jpress,
                                      Real* __restrict__ tau)
{
    const int64_t ncl = ncol * nlay;
    const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= ncl * k.nbnd) return;
    const int64_t idx = tid % ncl;
    const int ibnd = int(tid / ncl);
    const int64_t icol = idx % ncol, ilay = idx / ncol;
    const int itropo = !tropo[idx];
    const int gs_b = k.band_lims_gpt[2 * ibnd] - 1, ge_b = k.band_lims_gpt[2 * ibnd + 1];
    const Real p = play[idx], t = tlay[idx];
    for (int gg = gs_b; gg < ge_b; ++gg)
    {
        const int iflav = k.gpoint_flavor[itropo + 2 * gg] - 1;
        CL<Real> c; FL<Real> f;
        load_state(k, idx, ncl, iflav, tropo, jtemp, jpress, jeta, col_mix, fminor, fmajor, c, f);
        Real tt = tau[idx + ncl * gg] + major_tau(k, c, f, gg);
        // minor intervals of the active set covering this g-point, in interval order
        const int s = itropo;
        for (int m = 0; m < k.nminor[s]; ++m)
        {
            const int gs = k.minor_limits[s][2 * m] - 1, ge = k.minor_limits[s][2 * m + 1];
            if (gg < gs || gg >= ge) continue;
            Real scaling = col_gas[idx + ncl * k.idx_minor[s][m]];
            if (k.dens[s][m])
            {
                scaling *= Real(0.01) * p / t;
                const int is = k.idx_scal[s][m];
                if (is > 0)
                {
                    const Real vmr_fact = Real(1) / col_gas[idx];
                    const Real dry_fact = Real(1) / (Real(1) + col_gas[idx + ncl * k.idx_h2o] * vmr_fact);
                    if (k.compl_[s][m])
                        scaling *= (Real(1) - col_gas[idx + ncl * is] * vmr_fact * dry_fact);
                    else
                        scaling *= col_gas[idx + ncl * is] * vmr_fact * dry_fact;
                }
            }
            const int mflav = k.gpoint_flavor[2 * gs + s] - 1;
            CL<Real> cm; FL<Real> fm;
            load_state(k, idx, ncl, mflav, tropo, jtemp, jpress, jeta, col_mix, fminor, fmajor, cm, fm);
            const int64_t se = k.ntemp, sc = int64_t(k.ntemp) * k.neta;
            const int64_t off = sc * (gg - gs + k.kstart[s][m] - 1);
            const Real* kb0 = k.kminor[s] + (cm.jtemp - 1) + se * (fm.jeta[0] - 1) + off;
            const Real* kb1 = k.kminor[s] + (cm.jtemp) + se * (fm.jeta[1] - 1) + off;
            const Real lt = fm.fminor[0][0] * kb0[0] + fm.fminor[0][1] * kb0[se] +
                            fm.fminor[1][0] * kb1[0] + fm.fminor[1][1] * kb1[se];
            tt += lt * scaling;
        }
        tau[idx + ncl * gg] = tt;
        (void)icol; (void)ilay;
    }
}

template<typename Real>
__global__ void tau_rayleigh_kernel(Kd<Real> k, int64_t ncol, int64_t nlay, const Real* __restrict__ col_dry,
                                    const Real* __restrict__ col_gas, const Real* __restrict__ fminor,
                                    const int* __restrict__ jeta, const Bool* __restrict__ tropo,
                                    const int* __restrict__ jtemp, Real* __restrict__ tau_ray)
{
    const int64_t ncl = ncol * nlay;
    const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= ncl * k.ngpt) return;
    const int64_t idx = tid % ncl;
    const int gg = int(tid / ncl);
    const int itropo = !tropo[idx];
    const int iflav = k.gpoint_flavor[itropo + 2 * gg] - 1;
    CL<Real> c; FL<Real> f;
    load_state<Real>(k, idx, ncl, iflav, tropo, jtemp, nullptr, jeta, nullptr, fminor, nullptr, c, f);
    const int64_t se = k.ntemp;
    const Real* kr = k.krayl + int64_t(itropo) * k.ntemp * k.neta * k.ngpt + int64_t(k.ntemp) * k.neta * gg;
    const Real* k0 = kr + (c.jtemp - 1) + se * (f.jeta[0] - 1);
    const Real* k1 = kr + (c.jtemp) + se * (f.jeta[1] - 1);
    const Real kloc = f.fminor[0][0] * k0[0] + f.fminor[0][1] * k0[se] + f.fminor[1][0] * k1[0] + f.fminor[1][1] * k1[se];
    tau_ray[idx + ncl * gg] = kloc * (col_gas[idx + ncl * k.idx_h2o] + col_dry[idx]);
}

template<typename Real>
__global__ void combine_kernel(int64_t n, const Real* __restrict__ tau_abs, const Real* __restrict__ tau_ray,
                               Real* __restrict__ tau, Real* __restrict__ ssa, Real* __restrict__ g)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Real tr = tau_ray[i];
    const Real tt = tau_abs[i] + tr;
    ssa[i] = (tt > Real(2 * prop_eps)) ? tr / tt : Real(0);
    tau[i] = tt;
    g[i] = Real(0);
}

template<typename Real>
__global__ void planck_kernel(Kd<Real> k, int64_t ncol, int64_t nlay, const Real* __restrict__ tlay,
                              const Real* __restrict__ tlev, const Real* __restrict__ tsfc, int sfc_lay,
                              const Real* __restrict__ fmajor, const int* __restrict__ jeta,
                              const Bool* __restrict__ tropo, const int* __restrict__ jtemp,
                              const int* __restrict__ jpress, Real* __restrict__ sfc_src,
                              Real* __restrict__ lay_src, Real* __restrict__ lev_inc,
                              Real* __restrict__ lev_dec, Real* __restrict__ sfc_jac)
{
    const int64_t ncl = ncol * nlay;
    const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= ncl * k.ngpt) return;
    const int64_t idx = tid % ncl;
    const int gg = int(tid / ncl);
    const int64_t icol = idx % ncol, ilay = idx / ncol;
    const int itropo = !tropo[idx];
    const int iflav = k.gpoint_flavor[itropo + 2 * gg] - 1;
    CL<Real> c; FL<Real> f;
    load_state<Real>(k, idx, ncl, iflav, tropo, jtemp, jpress, jeta, nullptr, nullptr, fmajor, c, f);
    int ibnd = 0;
    while (gg >= k.band_lims_gpt[2 * ibnd + 1]) ++ibnd;
    const Real* tab = k.totplnk + int64_t(ibnd) * k.nplanck;
    const Real pf = planck_fraction(k, c, f, gg);
    lay_src[tid] = pf * interpolate1D(tlay[idx], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
    lev_inc[tid] = pf * interpolate1D(tlev[icol + ncol * (ilay + 1)], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
    lev_dec[tid] = pf * interpolate1D(tlev[idx], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
    if (ilay == sfc_lay - 1)
    {
        const Real s1 = interpolate1D(tsfc[icol], k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        const Real s2 = interpolate1D(tsfc[icol] + Real(1), k.temp_ref_min, k.totplnk_delta, k.nplanck, tab);
        sfc_src[icol + ncol * gg] = pf * s1;
        sfc_jac[icol + ncol * gg] = pf * (s2 - s1);
    }
}

// get_col_dry (src/Gas_optics_rrtmgp.cpp:764-792).
template<typename Real>
__global__ void col_dry_kernel(int64_t ncol, int64_t nlay, const Real* __restrict__ h2o, int64_t scol,
                               int64_t slay, const Real* __restrict__ plev, Real* __restrict__ col_dry)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ncol * nlay) return;
    const int64_t icol = i % ncol, ilay = i / ncol;
    const Real g0 = Real(9.80665), avogad = Real(6.02214076e23), m_dry = Real(0.028964), m_h2o = Real(0.018016);
    const Real v = h2o[icol * scol + ilay * slay];
    const Real dp = fabs(plev[i] - plev[i + ncol]);
    const Real m_air = (m_dry + m_h2o * v) / (Real(1) + v);
    Real cdry = Real(10) * dp * avogad / (Real(1000) * m_air * Real(100) * g0);
    cdry /= (Real(1) + v);
    col_dry[i] = cdry;
}

template<typename Real>
__global__ void fill_col_gas_kernel(int64_t ncol, int64_t nlay, int ngas, Gases<Real> g,
                                    const Real* __restrict__ col_dry, Real* __restrict__ col_gas)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ncl = ncol * nlay;
    if (i >= ncl) return;
    const int64_t icol = i % ncol, ilay = i / ncol;
    const Real cd = col_dry[i];
    col_gas[i] = cd;
    for (int ig = 1; ig <= ngas; ++ig) col_gas[i + ncl * ig] = col_of(g, ig, icol, ilay, cd);
}

// ---------------------------------------------------------------------------------------------
// Host launchers.

int check_kd(const rrtmgp_kdist_dev* kd, bool need_planck, bool need_rayleigh)
{
    if (!kd) return arg_error("k-distribution is NULL");
    if (kd->ngas < 1 || kd->ngas > RRTMGP_MAX_GAS) return arg_error("ngas out of range");
    if (!kd->kmajor || !kd->flavor || !kd->gpoint_flavor || !kd->band_lims_gpt)
        return arg_error("k-distribution tables are missing");
    if (need_planck && (!kd->totplnk || !kd->planck_frac || kd->nplanck < 2))
        return arg_error("k-distribution has no Planck tables (longwave entry on a shortwave kdist)");
    if (need_rayleigh && !kd->krayl)
        return arg_error("k-distribution has no Rayleigh tables (shortwave entry on a longwave kdist)");
    if (kd->idx_h2o < 1 || kd->idx_h2o > kd->ngas) return arg_error("idx_h2o cannot be found");
    return RRTMGP_OK;
}

// The fused kernels read the band-packed tables of kdist_pack.
int check_kd_fused(const rrtmgp_kdist_dev* kd, bool lw)
{
    RRTMGP_TRY(check_kd(kd, lw, !lw));
    if (!kd->go_bands || !kd->kmajor_p || !kd->kminor_p || (lw && !kd->planck_frac_p) || (!lw && !kd->krayl_p))
        return arg_error("k-distribution has no band-packed tables (call rrtmgp_hip_kdist_pack first)");
    if (kd->max_band_gpts < 1 || kd->max_band_gpts > kd->ngpt) return arg_error("max_band_gpts out of range");
    return RRTMGP_OK;
}

int check_go_grid(const GoGrid& g)
{
    if (g.ntl >= (int64_t(1) << 31))
        return arg_error("fused gas optics: too many (column tile, layer) units for one launch; "
                         "split the columns into blocks");
    return RRTMGP_OK;
}

// ---------------------------------------------------------------------------------------------
// Band-packed tables (layout described above the fused kernels), built on the host from the
// reference-layout tables of the k-distribution.

struct GoPack
{
    void* mem[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    ~GoPack()
    {
        for (void* p : mem)
            if (p) (void)hipFree(p);
    }
};

template<typename T>
int pack_d2h(std::vector<T>& out, const void* src, int64_t n, const char* what)
{
    out.assign(size_t(std::max<int64_t>(n, 0)), T());
    if (n <= 0) return RRTMGP_OK;
    if (!src) return arg_error(std::string("kdist_pack: ") + what + " is NULL");
    RRTMGP_HIP(hipMemcpy(out.data(), src, size_t(n) * sizeof(T), hipMemcpyDeviceToHost));
    return RRTMGP_OK;
}

template<typename T>
int pack_h2d(void** dst, const std::vector<T>& v)
{
    RRTMGP_HIP(hipMalloc(dst, std::max<size_t>(v.size(), 1) * sizeof(T)));
    if (!v.empty()) RRTMGP_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return RRTMGP_OK;
}

template<typename Real>
int kdist_pack(rrtmgp_kdist_dev* kd, void** pack_out)
{
    if (!kd || !pack_out) return arg_error("kdist_pack: NULL argument");
    *pack_out = nullptr;
    const bool lw = kd->planck_frac != nullptr;
    RRTMGP_TRY(check_kd(kd, lw, !lw));
    const int64_t nt = kd->ntemp, ne = kd->neta, np1 = kd->npres + 1, ngpt = kd->ngpt, nb = kd->nbnd;
    if (nt < 2 || ne < 2 || np1 < 3 || ngpt < 1 || nb < 1) return arg_error("kdist_pack: bad table dimensions");
    const int64_t nmin[2] = {kd->nminor_lower, kd->nminor_upper};
    const int64_t ncon[2] = {kd->ncontrib_lower, kd->ncontrib_upper};
    const void* kmin_src[2] = {kd->kminor_lower, kd->kminor_upper};
    const void* lim_src[2] = {kd->minor_limits_gpt_lower, kd->minor_limits_gpt_upper};
    const void* idx_src[2] = {kd->idx_minor_lower, kd->idx_minor_upper};
    const void* scal_src[2] = {kd->idx_minor_scaling_lower, kd->idx_minor_scaling_upper};
    const void* kst_src[2] = {kd->kminor_start_lower, kd->kminor_start_upper};
    const void* dens_src[2] = {kd->minor_scales_with_density_lower, kd->minor_scales_with_density_upper};
    const void* cmpl_src[2] = {kd->scale_by_complement_lower, kd->scale_by_complement_upper};

    std::vector<Real> kmaj, pfr, kray, kmin[2];
    std::vector<int> lims, gflav, flav, mlim[2], idxm[2], idxs[2], kst[2];
    std::vector<Bool> dens[2], cmpl[2];
    const int64_t nmaj = nt * ne * np1 * ngpt;
    RRTMGP_TRY(pack_d2h(kmaj, kd->kmajor, nmaj, "kmajor"));
    if (lw) RRTMGP_TRY(pack_d2h(pfr, kd->planck_frac, nmaj, "planck_frac"));
    else RRTMGP_TRY(pack_d2h(kray, kd->krayl, nt * ne * ngpt * 2, "krayl"));
    RRTMGP_TRY(pack_d2h(lims, kd->band_lims_gpt, 2 * nb, "band_lims_gpt"));
    RRTMGP_TRY(pack_d2h(gflav, kd->gpoint_flavor, 2 * ngpt, "gpoint_flavor"));
    RRTMGP_TRY(pack_d2h(flav, kd->flavor, 2 * kd->nflav, "flavor"));
    for (int s = 0; s < 2; ++s)
    {
        if (nmin[s] < 0 || ncon[s] < 0) return arg_error("kdist_pack: negative minor dimensions");
        RRTMGP_TRY(pack_d2h(kmin[s], kmin_src[s], nt * ne * ncon[s], "kminor"));
        RRTMGP_TRY(pack_d2h(mlim[s], lim_src[s], 2 * nmin[s], "minor_limits_gpt"));
        RRTMGP_TRY(pack_d2h(idxm[s], idx_src[s], nmin[s], "idx_minor"));
        RRTMGP_TRY(pack_d2h(idxs[s], scal_src[s], nmin[s], "idx_minor_scaling"));
        RRTMGP_TRY(pack_d2h(kst[s], kst_src[s], nmin[s], "kminor_start"));
        RRTMGP_TRY(pack_d2h(dens[s], dens_src[s], nmin[s], "minor_scales_with_density"));
        RRTMGP_TRY(pack_d2h(cmpl[s], cmpl_src[s], nmin[s], "scale_by_complement"));
    }

    auto gas_ok = [&](int gas) { return gas >= 0 && gas <= kd->ngas; };
    std::vector<GoBand> bands(static_cast<size_t>(nb));
    std::vector<int> minor_of(size_t(nb) * 2 * GO_MAXM, -1);
    int64_t koff = 0, roff = 0, moff = 0;
    int max_gpts = 0;
    for (int64_t b = 0; b < nb; ++b)
    {
        GoBand& B = bands[b];
        B = GoBand{};
        const int g0 = lims[2 * b] - 1, g1 = lims[2 * b + 1];        // g-points [g0, g1)
        if (g0 < 0 || g1 > ngpt || g1 <= g0) return arg_error("kdist_pack: bad band_lims_gpt");
        B.g0 = g0; B.ng = g1 - g0; B.gp = (B.ng + 3) / 4 * 4;
        max_gpts = std::max(max_gpts, B.ng);
        for (int s = 0; s < 2; ++s)
        {
            const int fl = gflav[2 * g0 + s];
            if (fl < 1 || fl > kd->nflav) return arg_error("kdist_pack: gpoint_flavor out of range");
            for (int gg = g0; gg < g1; ++gg)
                if (gflav[2 * gg + s] != fl) return arg_error("kdist_pack: the g-points of a band must share one flavor");
            B.gas[s][0] = flav[2 * (fl - 1)];
            B.gas[s][1] = flav[2 * (fl - 1) + 1];
            if (!gas_ok(B.gas[s][0]) || !gas_ok(B.gas[s][1])) return arg_error("kdist_pack: flavor gas out of range");
            int q = 0;
            for (int64_t m = 0; m < nmin[s]; ++m)
            {
                const int ms = mlim[s][2 * m] - 1, me = mlim[s][2 * m + 1] - 1;   // inclusive
                if (!(ms <= g1 - 1 && me >= g0)) continue;
                if (ms < 0 || me >= ngpt || me < ms) return arg_error("kdist_pack: bad minor_limits_gpt");
                if (q == GO_MAXM)
                    return arg_error("kdist_pack: more than " + std::to_string(GO_MAXM) + " minor intervals in a band");
                if (gflav[2 * ms + s] != fl)
                    return arg_error("kdist_pack: a minor interval does not have its band's flavor");
                if (!gas_ok(idxm[s][m]) || idxs[s][m] > kd->ngas) return arg_error("kdist_pack: minor gas out of range");
                if (kst[s][m] < 1 || kst[s][m] - 1 + (me - ms) >= ncon[s])
                    return arg_error("kdist_pack: kminor_start out of range");
                B.minor[s][q] = GoMinor{idxm[s][m], dens[s][m] ? 1 : 0, idxs[s][m], cmpl[s][m] ? 1 : 0};
                minor_of[(b * 2 + s) * GO_MAXM + q] = int(m);
                ++q;
            }
            B.nq[s] = q;
        }
        B.nslot = std::max(B.nq[0], B.nq[1]);
        B.koff = int(koff); B.roff = int(roff); B.moff = int(moff);
        koff += nt * ne * np1 * B.gp;
        roff += 2 * nt * ne * B.gp;
        moff += 2 * nt * ne * B.nslot * B.gp;
        const int64_t lim = (int64_t(1) << 32) / int64_t(sizeof(Real));    // 32-bit byte offsets
        if (koff >= lim || moff >= lim || roff >= lim)
            return arg_error("kdist_pack: tables too large for 32-bit offsets");
    }

    std::vector<Real> kmaj_p(size_t(koff), Real(0)), pfr_p(lw ? size_t(koff) : 0, Real(0));
    std::vector<Real> kray_p(lw ? 0 : size_t(roff), Real(0)), kmin_p(size_t(moff), Real(0));
    for (int64_t b = 0; b < nb; ++b)
    {
        const GoBand& B = bands[b];
        for (int64_t t = 0; t < nt; ++t)
            for (int64_t e = 0; e < ne; ++e)
            {
                for (int64_t p = 0; p < np1; ++p)
                    for (int gl = 0; gl < B.ng; ++gl)
                    {
                        const size_t dst = size_t(B.koff + ((t * ne + e) * np1 + p) * B.gp + gl);
                        const size_t src = size_t(t + nt * (e + ne * (p + np1 * (B.g0 + gl))));
                        kmaj_p[dst] = kmaj[src];
                        if (lw) pfr_p[dst] = pfr[src];
                    }
                for (int s = 0; s < 2; ++s)
                {
                    if (!lw)
                        for (int gl = 0; gl < B.ng; ++gl)
                            kray_p[size_t(B.roff + ((s * nt + t) * ne + e) * B.gp + gl)] =
                                    kray[size_t(t + nt * (e + ne * ((B.g0 + gl) + ngpt * s)))];
                    for (int q = 0; q < B.nq[s]; ++q)
                    {
                        const int m = minor_of[(b * 2 + s) * GO_MAXM + q];
                        const int ms = mlim[s][2 * m] - 1, me = mlim[s][2 * m + 1] - 1;
                        for (int gl = 0; gl < B.ng; ++gl)
                        {
                            const int gg = B.g0 + gl;
                            if (gg < ms || gg > me) continue;
                            const int64_t ic = kst[s][m] - 1 + (gg - ms);
                            kmin_p[size_t(B.moff + (((s * nt + t) * ne + e) * B.nslot + q) * B.gp + gl)] =
                                    kmin[s][size_t(t + nt * (e + ne * ic))];
                        }
                    }
                }
            }
    }

    auto pk = std::make_unique<GoPack>();
    RRTMGP_TRY(pack_h2d(&pk->mem[0], bands));
    RRTMGP_TRY(pack_h2d(&pk->mem[1], kmaj_p));
    RRTMGP_TRY(pack_h2d(&pk->mem[2], kmin_p));
    if (lw) RRTMGP_TRY(pack_h2d(&pk->mem[3], pfr_p));
    else RRTMGP_TRY(pack_h2d(&pk->mem[4], kray_p));
    kd->go_bands = pk->mem[0];
    kd->kmajor_p = pk->mem[1];
    kd->kminor_p = pk->mem[2];
    kd->planck_frac_p = pk->mem[3];
    kd->krayl_p = pk->mem[4];
    kd->max_band_gpts = max_gpts;
    *pack_out = pk.release();
    return RRTMGP_OK;
}

template<typename Real>
int gas_optics_lw(const rrtmgp_kdist_dev* kd, int64_t b0, int64_t nb, int64_t ncol, int64_t nlay, const Real* play,
                  const Real* tlay, const Real* tlev, const Real* tsfc, const Real* col_dry,
                  const rrtmgp_gas_vmr* gases, int sfc_lay, const Real* cloud_tau, Real* tau,
                  Real* lay_src, Real* lev_inc, Real* lev_dec, Real* sfc_src, Real* sfc_jac,
                  void* stream, const int* tiles = nullptr, int64_t ntiles = 0, Real* bnd_lay = nullptr,
                  Real* bnd_lev = nullptr)
{
    RRTMGP_TRY(check_kd_fused(kd, true));
    if (ncol <= 0 || nlay <= 0) return arg_error("ncol and nlay must be positive");
    if (sfc_lay != 1 && sfc_lay != nlay) return arg_error("sfc_lay must be 1 or nlay");
    const bool factored = bnd_lay != nullptr;
    if (factored != (bnd_lev != nullptr)) return arg_error("gas_optics_lw: band Planck arrays go together");
    if (!play || !tlay || !tlev || !tsfc || !col_dry || !tau || !lay_src || !sfc_src || !sfc_jac ||
        (!factored && (!lev_inc || !lev_dec)))
        return arg_error("gas_optics_lw: NULL array");
    Gases<Real> g;
    RRTMGP_TRY(make_gases(gases, kd->ngas, ncol, nlay, g));
    const Kd<Real> k = make_kd<Real>(*kd);
    if (b0 < 0 || nb < 1 || b0 + nb > kd->nbnd) return arg_error("gas_optics_lw: band range outside the k-distribution");
    if (ntiles < 0 || ntiles > ((ncol + GO_COLS - 1) / GO_COLS) * nlay || (ntiles > 0 && !tiles))
        return arg_error("fused gas optics: bad tile list");
    if (tiles && ntiles == 0) return RRTMGP_OK;                 // nothing listed
    const GoGrid grid = make_go_grid(ncol, nlay, b0, nb, kd->max_band_gpts, tiles, ntiles);
    RRTMGP_TRY(check_go_grid(grid));
    RRTMGP_TIMED("gas_optics_lw_kernel", as_stream(stream));
    const size_t dyn_lw = size_t(go_in_rows<true>(kd->ngas)) * GO_COLS * sizeof(Real);   // phase 1's input rows
    hipLaunchKernelGGL(gas_optics_lw_kernel<Real>, dim3(xcd_grid(grid.ntl)), dim3(GO_THREADS), dyn_lw,
                       as_stream(stream), k, g, ncol, nlay, grid, play, tlay, tlev, tsfc, col_dry,
                       sfc_lay, cloud_tau, tau, lay_src, lev_inc, lev_dec, sfc_src, sfc_jac, bnd_lay, bnd_lev);
    RRTMGP_LAUNCHED("gas_optics_lw_kernel");
    return RRTMGP_OK;
}

template<typename Real>
int gas_optics_sw(const rrtmgp_kdist_dev* kd, int64_t b0, int64_t nb, int64_t ncol, int64_t nlay, const Real* play,
                  const Real* tlay, const Real* col_dry, const rrtmgp_gas_vmr* gases,
                  const Real* ctau, const Real* cssa, const Real* cg, Real* tau, Real* ssa, Real* g,
                  void* stream, const int* tiles = nullptr, int64_t ntiles = 0, bool coalb = false)
{
    RRTMGP_TRY(check_kd_fused(kd, false));
    if (ncol <= 0 || nlay <= 0) return arg_error("ncol and nlay must be positive");
    if (!play || !tlay || !col_dry || !tau || !ssa || !g) return arg_error("gas_optics_sw: NULL array");
    if (ctau && (!cssa || !cg)) return arg_error("cloud ssa/g missing");
    Gases<Real> gs;
    RRTMGP_TRY(make_gases(gases, kd->ngas, ncol, nlay, gs));
    const Kd<Real> k = make_kd<Real>(*kd);
    if (b0 < 0 || nb < 1 || b0 + nb > kd->nbnd) return arg_error("gas_optics_sw: band range outside the k-distribution");
    if (ntiles < 0 || ntiles > ((ncol + GO_COLS - 1) / GO_COLS) * nlay || (ntiles > 0 && !tiles))
        return arg_error("fused gas optics: bad tile list");
    if (tiles && ntiles == 0) return RRTMGP_OK;                 // nothing listed
    const GoGrid grid = make_go_grid(ncol, nlay, b0, nb, kd->max_band_gpts, tiles, ntiles);
    RRTMGP_TRY(check_go_grid(grid));
    RRTMGP_TIMED("gas_optics_sw_kernel", as_stream(stream));
    const size_t dyn_sw = size_t(go_in_rows<false>(kd->ngas)) * GO_COLS * sizeof(Real);
    hipLaunchKernelGGL(gas_optics_sw_kernel<Real>, dim3(xcd_grid(grid.ntl)), dim3(GO_THREADS), dyn_sw,
                       as_stream(stream), k, gs, ncol, nlay, grid, play, tlay, col_dry, ctau, cssa, cg,
                       tau, ssa, g, coalb ? 1 : 0);
    RRTMGP_LAUNCHED("gas_optics_sw_kernel");
    return RRTMGP_OK;
}

template<typename Real>
int gas_optics_adjoint(const rrtmgp_kdist_dev* kd, bool lw, int64_t band, int64_t ncol, int64_t nlay,
                       const Real* play, const Real* tlay, const Real* tlev, const Real* tsfc, const Real* col_dry,
                       const rrtmgp_gas_vmr* gases, int sfc_lay, const Real* ctau, const Real* cssa, const Real* cg,
                       const Real* ga, const Real* gb, const Real* gc, const Real* gd, const Real* gs, Real* d_col,
                       Real* d_tlay, Real* d_lnp, Real* d_tlev_lo, Real* d_tlev_hi, Real* d_tsfc, void* stream)
{
    RRTMGP_TRY(check_kd_fused(kd, lw));
    if (ncol <= 0 || nlay <= 0) return arg_error("gas_optics_adjoint: ncol and nlay must be positive");
    if (band < 0 || band >= kd->nbnd) return arg_error("gas_optics_adjoint: band outside the k-distribution");
    if (!play || !tlay || !col_dry || !ga || !gb || !gc || !d_col || !d_tlay || !d_lnp)
        return arg_error("gas_optics_adjoint: NULL array");
    if (lw && (!tlev || !tsfc || !gd || !gs || !d_tlev_lo || !d_tlev_hi || !d_tsfc || (sfc_lay != 1 && sfc_lay != nlay)))
        return arg_error("gas_optics_adjoint: longwave arrays missing or sfc_lay not 1 or nlay");
    if (!lw && ctau && (!cssa || !cg)) return arg_error("gas_optics_adjoint: cloud ssa/g missing");
    Gases<Real> g;
    RRTMGP_TRY(make_gases(gases, kd->ngas, ncol, nlay, g));
    const Kd<Real> k = make_kd<Real>(*kd);
    const int64_t nunits = ((ncol + GO_COLS - 1) / GO_COLS) * nlay;
    if (nunits >= (int64_t(1) << 31)) return arg_error("gas_optics_adjoint: too many column tiles x layers");
    RRTMGP_TIMED("gas_optics_adjoint_kernel", as_stream(stream));
    if (lw)
        hipLaunchKernelGGL((gas_optics_adjoint_kernel<Real, true>), dim3(xcd_grid(nunits)), dim3(GO_THREADS), 0,
                           as_stream(stream), k, g, ncol, nlay, int(band), play, tlay, tlev, tsfc, col_dry, sfc_lay,
                           nullptr, nullptr, nullptr, ga, gb, gc, gd, gs, d_col, d_tlay, d_lnp, d_tlev_lo,
                           d_tlev_hi, d_tsfc);
    else
        hipLaunchKernelGGL((gas_optics_adjoint_kernel<Real, false>), dim3(xcd_grid(nunits)), dim3(GO_THREADS), 0,
                           as_stream(stream), k, g, ncol, nlay, int(band), play, tlay, nullptr, nullptr, col_dry, 1,
                           ctau, cssa, cg, ga, gb, gc, nullptr, nullptr, d_col, d_tlay, d_lnp, nullptr, nullptr,
                           nullptr);
    RRTMGP_LAUNCHED("gas_optics_adjoint_kernel");
    return RRTMGP_OK;
}

// Adjoint of get_col_dry (col_dry_kernel): col_dry = K |dp| / (m_dry + m_h2o v) exactly, so
// d/dv = -col_dry m_h2o / (m_dry + m_h2o v) and d/dp_lev = +-K / (m_dry + m_h2o v).  Adds into
// d_h2o (ncol, nlay) and d_plev (ncol, nlev), one thread per column.
template<typename Real>
__global__ void col_dry_adjoint_kernel(int64_t ncol, int64_t nlay, const Real* __restrict__ h2o, int64_t scol,
                                       int64_t slay, const Real* __restrict__ plev,
                                       const Real* __restrict__ g_col_dry, Real* __restrict__ d_h2o,
                                       Real* __restrict__ d_plev)
{
    const int64_t icol = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (icol >= ncol) return;
    const Real g0 = Real(9.80665), avogad = Real(6.02214076e23), m_dry = Real(0.028964), m_h2o = Real(0.018016);
    const Real Kc = Real(10) * avogad / (Real(1000) * Real(100) * g0);
    for (int64_t ilay = 0; ilay < nlay; ++ilay)
    {
        const int64_t i = icol + ncol * ilay;
        const Real v = h2o[icol * scol + ilay * slay];
        const Real m = m_dry + m_h2o * v;
        const Real dp = plev[i] - plev[i + ncol];
        const Real gc = g_col_dry[i];
        d_h2o[i] += -gc * Kc * fabs(dp) * m_h2o / (m * m);
        const Real sgn = dp >= Real(0) ? Real(1) : Real(-1);
        d_plev[i] += gc * Kc * sgn / m;
        d_plev[i + ncol] -= gc * Kc * sgn / m;
    }
}

template<typename Real>
int col_dry_adjoint(int64_t ncol, int64_t nlay, const Real* h2o, int64_t d1, int64_t d2, const Real* plev,
                    const Real* g_col_dry, Real* d_h2o, Real* d_plev, void* stream)
{
    if (ncol <= 0 || nlay <= 0 || !h2o || !plev || !g_col_dry || !d_h2o || !d_plev)
        return arg_error("col_dry_adjoint: bad arguments");
    if (!((d1 == 1 || d1 == ncol) && (d2 == 1 || d2 == nlay))) return arg_error("col_dry_adjoint: bad h2o dims");
    const int64_t scol = d1 == 1 ? 0 : 1, slay = d2 == 1 ? 0 : d1;
    RRTMGP_TIMED("col_dry_adjoint_kernel", as_stream(stream));
    hipLaunchKernelGGL(col_dry_adjoint_kernel<Real>, dim3(grid1(ncol, 256)), dim3(256), 0, as_stream(stream),
                       ncol, nlay, h2o, scol, slay, plev, g_col_dry, d_h2o, d_plev);
    RRTMGP_LAUNCHED("col_dry_adjoint_kernel");
    return RRTMGP_OK;
}
}  // namespace rrtmgp_hip

using namespace rrtmgp_hip;

#define RRTMGP_DEF_GAS_OPTICS(S, T)                                                               \
extern "C" int rrtmgp_hip_gas_optics_lw_##S(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay, \
    const T* play, const T* tlay, const T* tlev, const T* tsfc, const T* col_dry,                   \
    const rrtmgp_gas_vmr* gases, int sfc_lay, const T* cloud_tau, T* tau, T* lay_src,              \
    T* lev_src_inc, T* lev_src_dec, T* sfc_src, T* sfc_src_jac, void* stream)                      \
{                                                                                                  \
    return gas_optics_lw<T>(kd, 0, kd ? kd->nbnd : 0, ncol, nlay, play, tlay, tlev, tsfc, col_dry, \
                            gases, sfc_lay, cloud_tau, tau, lay_src, lev_src_inc, lev_src_dec,     \
                            sfc_src, sfc_src_jac, stream);                                         \
}                                                                                                  \
extern "C" int rrtmgp_hip_gas_optics_lw_factored_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,     \
    int64_t nlay, const int* tiles, int64_t ntiles, const T* play, const T* tlay, const T* tlev,    \
    const T* tsfc, const T* col_dry, const rrtmgp_gas_vmr* gases, int sfc_lay, const T* cloud_tau,  \
    T* tau, T* planck_frac, T* band_lay_src, T* band_lev_src, T* sfc_src, T* sfc_src_jac,          \
    void* stream)                                                                                   \
{                                                                                                  \
    if (!band_lay_src || !band_lev_src) return arg_error("gas_optics_lw_factored: NULL band array"); \
    if (!tiles && ntiles != 0) return arg_error("gas_optics_lw_factored: tile count without a list"); \
    return gas_optics_lw<T>(kd, 0, kd ? kd->nbnd : 0, ncol, nlay, play, tlay, tlev, tsfc, col_dry, \
                            gases, sfc_lay, cloud_tau, tau, planck_frac, nullptr, nullptr,         \
                            sfc_src, sfc_src_jac, stream, tiles, ntiles, band_lay_src,            \
                            band_lev_src);                                                         \
}                                                                                                  \
extern "C" int rrtmgp_hip_gas_optics_lw_bands_##S(const rrtmgp_kdist_dev* kd, int64_t band_start,  \
    int64_t band_count, int64_t ncol, int64_t nlay, const T* play, const T* tlay, const T* tlev,    \
    const T* tsfc, const T* col_dry, const rrtmgp_gas_vmr* gases, int sfc_lay, const T* cloud_tau,  \
    T* tau, T* lay_src, T* lev_src_inc, T* lev_src_dec, T* sfc_src, T* sfc_src_jac, void* stream)   \
{                                                                                                  \
    return gas_optics_lw<T>(kd, band_start, band_count, ncol, nlay, play, tlay, tlev, tsfc,       \
                            col_dry, gases, sfc_lay, cloud_tau, tau, lay_src, lev_src_inc,         \
                            lev_src_dec, sfc_src, sfc_src_jac, stream);                            \
}                                                                                                  \
extern "C" int rrtmgp_hip_gas_optics_sw_##S(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay, \
    const T* play, const T* tlay, const T* col_dry, const rrtmgp_gas_vmr* gases,                   \
    const T* cloud_tau, const T* cloud_ssa, const T* cloud_g, T* tau, T* ssa, T* g, void* stream)  \
{                                                                                                  \
    return gas_optics_sw<T>(kd, 0, kd ? kd->nbnd : 0, ncol, nlay, play, tlay, col_dry, gases,      \
                            cloud_tau, cloud_ssa, cloud_g, tau, ssa, g, stream);                   \
}                                                                                                  \
extern "C" int rrtmgp_hip_gas_optics_sw_coalbedo_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,   \
    int64_t nlay, const T* play, const T* tlay, const T* col_dry, const rrtmgp_gas_vmr* gases,      \
    const T* cloud_tau, const T* cloud_coalbedo, const T* cloud_g, T* tau, T* coalbedo, T* g,       \
    void* stream)                                                                                   \
{                                                                                                  \
    return gas_optics_sw<T>(kd, 0, kd ? kd->nbnd : 0, ncol, nlay, play, tlay, col_dry, gases,      \
                            cloud_tau, cloud_coalbedo, cloud_g, tau, coalbedo, g, stream, nullptr, \
                            0, true);                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_gas_optics_sw_bands_##S(const rrtmgp_kdist_dev* kd, int64_t band_start,  \
    int64_t band_count, int64_t ncol, int64_t nlay, const T* play, const T* tlay, const T* col_dry, \
    const rrtmgp_gas_vmr* gases, const T* cloud_tau, const T* cloud_ssa, const T* cloud_g, T* tau,  \
    T* ssa, T* g, void* stream)                                                                     \
{                                                                                                  \
    return gas_optics_sw<T>(kd, band_start, band_count, ncol, nlay, play, tlay, col_dry, gases,    \
                            cloud_tau, cloud_ssa, cloud_g, tau, ssa, g, stream);                   \
}                                                                                                  \
extern "C" int rrtmgp_hip_gas_optics_lw_tiles_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,      \
    int64_t nlay, const int* tiles, int64_t ntiles, const T* play, const T* tlay, const T* tlev,    \
    const T* tsfc, const T* col_dry, const rrtmgp_gas_vmr* gases, int sfc_lay, const T* cloud_tau,  \
    T* tau, T* lay_src, T* lev_src_inc, T* lev_src_dec, T* sfc_src, T* sfc_src_jac, void* stream)   \
{                                                                                                  \
    if (!tiles) return arg_error("gas_optics_lw_tiles: NULL tile list");                          \
    return gas_optics_lw<T>(kd, 0, kd ? kd->nbnd : 0, ncol, nlay, play, tlay, tlev, tsfc, col_dry, \
                            gases, sfc_lay, cloud_tau, tau, lay_src, lev_src_inc, lev_src_dec,     \
                            sfc_src, sfc_src_jac, stream, tiles, ntiles);                          \
}                                                                                                  \
extern "C" int rrtmgp_hip_gas_optics_sw_tiles_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,      \
    int64_t nlay, const int* tiles, int64_t ntiles, const T* play, const T* tlay, const T* col_dry, \
    const rrtmgp_gas_vmr* gases, const T* cloud_tau, const T* cloud_ssa, const T* cloud_g, T* tau,  \
    T* ssa, T* g, void* stream)                                                                     \
{                                                                                                  \
    if (!tiles) return arg_error("gas_optics_sw_tiles: NULL tile list");                          \
    return gas_optics_sw<T>(kd, 0, kd ? kd->nbnd : 0, ncol, nlay, play, tlay, col_dry, gases,      \
                            cloud_tau, cloud_ssa, cloud_g, tau, ssa, g, stream, tiles, ntiles);    \
}                                                                                                  \
extern "C" int rrtmgp_hip_interpolation_##S(const rrtmgp_kdist_dev* kd, int64_t ncol, int64_t nlay, \
    const T* play, const T* tlay, const T* col_gas, int* jtemp, T* fmajor, T* fminor, T* col_mix,  \
    rrtmgp_bool* tropo, int* jeta, int* jpress, void* stream)                                      \
{                                                                                                  \
    RRTMGP_TRY(check_kd(kd, false, false));                                                       \
    const Kd<T> k = make_kd<T>(*kd);                                                               \
    const int64_t n = ncol * nlay;                                                                 \
    RRTMGP_TIMED("interpolation_kernel", as_stream(stream));                                       \
    hipLaunchKernelGGL(interpolation_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                \
                       as_stream(stream), k, ncol, nlay, play, tlay, col_gas, jtemp, fmajor,       \
                       fminor, col_mix, tropo, jeta, jpress);                                      \
    RRTMGP_LAUNCHED("interpolation_kernel");                                                      \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_compute_tau_absorption_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,   \
    int64_t nlay, const rrtmgp_bool* tropo, const T* col_mix, const T* fmajor, const T* fminor,    \
    const T* play, const T* tlay, const T* col_gas, const int* jeta, const int* jtemp,             \
    const int* jpress, T* tau, void* stream)                                                       \
{                                                                                                  \
    RRTMGP_TRY(check_kd(kd, false, false));                                                       \
    const Kd<T> k = make_kd<T>(*kd);                                                               \
    const int64_t n = ncol * nlay * kd->nbnd;                                                      \
    RRTMGP_TIMED("tau_absorption_kernel", as_stream(stream));                                      \
    hipLaunchKernelGGL(tau_absorption_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,               \
                       as_stream(stream), k, ncol, nlay, tropo, col_mix, fmajor, fminor, play,     \
                       tlay, col_gas, jeta, jtemp, jpress, tau);                                   \
    RRTMGP_LAUNCHED("tau_absorption_kernel");                                                     \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_compute_tau_rayleigh_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,     \
    int64_t nlay, const T* col_dry, const T* col_gas, const T* fminor, const int* jeta,            \
    const rrtmgp_bool* tropo, const int* jtemp, T* tau_rayleigh, void* stream)                     \
{                                                                                                  \
    RRTMGP_TRY(check_kd(kd, false, true));                                                        \
    const Kd<T> k = make_kd<T>(*kd);                                                               \
    const int64_t n = ncol * nlay * kd->ngpt;                                                      \
    RRTMGP_TIMED("tau_rayleigh_kernel", as_stream(stream));                                        \
    hipLaunchKernelGGL(tau_rayleigh_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                 \
                       as_stream(stream), k, ncol, nlay, col_dry, col_gas, fminor, jeta, tropo,    \
                       jtemp, tau_rayleigh);                                                       \
    RRTMGP_LAUNCHED("tau_rayleigh_kernel");                                                       \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_combine_abs_and_rayleigh_##S(int64_t ncol, int64_t nlay, int64_t ngpt,  \
    const T* tau_abs, const T* tau_rayleigh, T* tau, T* ssa, T* g, void* stream)                   \
{                                                                                                  \
    const int64_t n = ncol * nlay * ngpt;                                                          \
    RRTMGP_TIMED("combine_kernel", as_stream(stream));                                             \
    hipLaunchKernelGGL(combine_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),   \
                       n, tau_abs, tau_rayleigh, tau, ssa, g);                                     \
    RRTMGP_LAUNCHED("combine_kernel");                                                            \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_compute_planck_source_##S(const rrtmgp_kdist_dev* kd, int64_t ncol,    \
    int64_t nlay, const T* tlay, const T* tlev, const T* tsfc, int sfc_lay, const T* fmajor,       \
    const int* jeta, const rrtmgp_bool* tropo, const int* jtemp, const int* jpress, T* sfc_src,    \
    T* lay_src, T* lev_src_inc, T* lev_src_dec, T* sfc_src_jac, void* stream)                      \
{                                                                                                  \
    RRTMGP_TRY(check_kd(kd, true, false));                                                        \
    const Kd<T> k = make_kd<T>(*kd);                                                               \
    const int64_t n = ncol * nlay * kd->ngpt;                                                      \
    RRTMGP_TIMED("planck_kernel", as_stream(stream));                                              \
    hipLaunchKernelGGL(planck_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),    \
                       k, ncol, nlay, tlay, tlev, tsfc, sfc_lay, fmajor, jeta, tropo, jtemp,       \
                       jpress, sfc_src, lay_src, lev_src_inc, lev_src_dec, sfc_src_jac);           \
    RRTMGP_LAUNCHED("planck_kernel");                                                             \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_get_col_dry_##S(int64_t ncol, int64_t nlay, const T* vmr_h2o,          \
    int64_t h2o_dim1, int64_t h2o_dim2, const T* plev, T* col_dry, void* stream)                   \
{                                                                                                  \
    const int64_t n = ncol * nlay;                                                                 \
    const int64_t scol = h2o_dim1 == 1 ? 0 : 1, slay = h2o_dim2 == 1 ? 0 : h2o_dim1;              \
    RRTMGP_TIMED("col_dry_kernel", as_stream(stream));                                             \
    hipLaunchKernelGGL(col_dry_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0, as_stream(stream),   \
                       ncol, nlay, vmr_h2o, scol, slay, plev, col_dry);                            \
    RRTMGP_LAUNCHED("col_dry_kernel");                                                            \
    return RRTMGP_OK;                                                                              \
}                                                                                                  \
extern "C" int rrtmgp_hip_fill_col_gas_##S(int64_t ncol, int64_t nlay, const rrtmgp_gas_vmr* gases,\
    const T* col_dry, T* col_gas, void* stream)                                                    \
{                                                                                                  \
    Gases<T> g;                                                                                    \
    RRTMGP_TRY(make_gases(gases, gases ? gases->ngas : 0, ncol, nlay, g));                        \
    const int64_t n = ncol * nlay;                                                                 \
    RRTMGP_TIMED("fill_col_gas_kernel", as_stream(stream));                                        \
    hipLaunchKernelGGL(fill_col_gas_kernel<T>, dim3(grid1(n, 256)), dim3(256), 0,                 \
                       as_stream(stream), ncol, nlay, int(gases->ngas), g, col_dry, col_gas);      \
    RRTMGP_LAUNCHED("fill_col_gas_kernel");                                                       \
    return RRTMGP_OK;                                                                              \
}

RRTMGP_DEF_GAS_OPTICS(f32, float)
RRTMGP_DEF_GAS_OPTICS(f64, double)

extern "C" int rrtmgp_hip_kdist_pack_f32(rrtmgp_kdist_dev* kd, void** pack) { return kdist_pack<float>(kd, pack); }
extern "C" int rrtmgp_hip_kdist_pack_f64(rrtmgp_kdist_dev* kd, void** pack) { return kdist_pack<double>(kd, pack); }
#define RRTMGP_DEF_GO_ADJOINT(S, T)                                                                 \
extern "C" int rrtmgp_hip_gas_optics_lw_adjoint_##S(const rrtmgp_kdist_dev* kd, int64_t band,       \
    int64_t ncol, int64_t nlay, const T* play, const T* tlay, const T* tlev, const T* tsfc,          \
    const T* col_dry, const rrtmgp_gas_vmr* gases, int sfc_lay, const T* g_tau, const T* g_lay_src,  \
    const T* g_lev_src_inc, const T* g_lev_src_dec, const T* g_sfc_src, T* d_col_gas, T* d_tlay,     \
    T* d_lnp_lay, T* d_tlev_lo, T* d_tlev_hi, T* d_tsfc, void* stream)                               \
{                                                                                                    \
    return gas_optics_adjoint<T>(kd, true, band, ncol, nlay, play, tlay, tlev, tsfc, col_dry, gases, \
                                 sfc_lay, nullptr, nullptr, nullptr, g_tau, g_lay_src, g_lev_src_inc,\
                                 g_lev_src_dec, g_sfc_src, d_col_gas, d_tlay, d_lnp_lay, d_tlev_lo,  \
                                 d_tlev_hi, d_tsfc, stream);                                         \
}                                                                                                    \
extern "C" int rrtmgp_hip_gas_optics_sw_adjoint_##S(const rrtmgp_kdist_dev* kd, int64_t band,       \
    int64_t ncol, int64_t nlay, const T* play, const T* tlay, const T* col_dry,                      \
    const rrtmgp_gas_vmr* gases, const T* cloud_tau, const T* cloud_ssa, const T* cloud_g,           \
    const T* g_tau, const T* g_ssa, const T* g_g, T* d_col_gas, T* d_tlay, T* d_lnp_lay, void* stream)\
{                                                                                                    \
    return gas_optics_adjoint<T>(kd, false, band, ncol, nlay, play, tlay, nullptr, nullptr, col_dry, \
                                 gases, 1, cloud_tau, cloud_ssa, cloud_g, g_tau, g_ssa, g_g, nullptr, \
                                 nullptr, d_col_gas, d_tlay, d_lnp_lay, nullptr, nullptr, nullptr,   \
                                 stream);                                                            \
}                                                                                                    \
extern "C" int rrtmgp_hip_get_col_dry_adjoint_##S(int64_t ncol, int64_t nlay, const T* vmr_h2o,     \
    int64_t h2o_dim1, int64_t h2o_dim2, const T* plev, const T* g_col_dry, T* d_vmr_h2o, T* d_plev,  \
    void* stream)                                                                                    \
{                                                                                                    \
    return col_dry_adjoint<T>(ncol, nlay, vmr_h2o, h2o_dim1, h2o_dim2, plev, g_col_dry, d_vmr_h2o,   \
                              d_plev, stream);                                                       \
}
RRTMGP_DEF_GO_ADJOINT(f32, float)
RRTMGP_DEF_GO_ADJOINT(f64, double)

extern "C" int rrtmgp_hip_kdist_pack_free(void* pack)
{
    delete static_cast<GoPack*>(pack);
    return RRTMGP_OK;
}
