// Coefficient sets by variable name, and the loaders of the k-distributions and cloud look-up
// tables (reference src_test/Radiation_solver.cpp:70-327, load_and_init_gas_optics /
// load_and_init_cloud_optics, which read NetCDF through include_test/Netcdf_interface.h).
//
// A bundle is a directory holding a coefficient file's variables under the reference's names:
// manifest.txt lists each dimension ("dim <name> <size>") and variable ("var <name> <type> <ndim>
// <dim names...>", type f8 f4 i4 i1 S1), and <name>.bin holds the values in C order over the
// NetCDF dimensions, little-endian.  rrtmgp_amd.netcdf converts a NetCDF-3 or NetCDF-4 coefficient
// file into one.  The accessors follow the reference's Netcdf_file (get_dimension_size,
// get_variable<T>(name, dims), variable_exists) so the loaders read like the reference's.
#ifndef RRTMGP_AMD_COEFFICIENT_BUNDLE_H
#define RRTMGP_AMD_COEFFICIENT_BUNDLE_H

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "Array.h"
#include "types.h"

class Gas_concs_gpu;
class Gas_optics_rrtmgp_gpu;
class Cloud_optics_gpu;

class Coefficient_bundle
{
    public:
        explicit Coefficient_bundle(const std::string& dir);

        int get_dimension_size(const std::string& name) const;
        bool variable_exists(const std::string& name) const;
        // Dimension names of a variable, slowest first (empty for a scalar).
        std::vector<std::string> get_variable_dimensions(const std::string& name) const;

        // All values of a variable, converted to T; dims (slowest first, as the NetCDF
        // dimensions) must match the stored shape.  An empty dims reads a scalar.
        template<typename T>
        std::vector<T> get_variable(const std::string& name, const std::vector<int>& dims) const;
        template<typename T>
        T get_variable(const std::string& name) const { return get_variable<T>(name, {}).at(0); }

        // A (n, string_len) character variable as n strings, trailing blanks removed.
        std::vector<std::string> get_strings(const std::string& name) const;

    private:
        struct Var { std::string type; std::vector<std::string> dims; };
        const Var& var(const std::string& name) const;
        std::vector<char> raw(const std::string& name, size_t elem_size, int64_t count) const;

        std::string dir;
        std::map<std::string, int> dims;
        std::map<std::string, Var> vars;
};

// Longwave or shortwave k-distribution of a bundle (src_test/Radiation_solver.cpp:70-284): a
// bundle with totplnk gives the longwave variant, otherwise the shortwave one.
std::unique_ptr<Gas_optics_rrtmgp_gpu> load_and_init_gas_optics(const Gas_concs_gpu& gas_concs,
                                                                const std::string& bundle_dir);

// Cloud look-up tables of a bundle (src_test/Radiation_solver.cpp:286-327).
std::unique_ptr<Cloud_optics_gpu> load_and_init_cloud_optics(const std::string& bundle_dir);

#endif
