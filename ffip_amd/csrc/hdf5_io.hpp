// HDF5 file access for the run_sim_json drop-in: the input file's float64
// datasets and the output file's groups (ffip/simulation.py:404-431 writes and
// reads them with h5py; src/simulation.cpp:151-157, 190-233 and
// src/dft.cpp:273-320 open and fill them with HighFive).
//
// libhdf5 (1.10 or later) is loaded at run time with dlopen, so the simulation
// library itself does not depend on it.  The search order is $FFIP_HDF5_LIBRARY,
// the system's libhdf5.so / libhdf5_serial.so, then /opt/conda/lib/libhdf5.so.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace ffip_amd {
namespace h5 {

// Path of the loaded libhdf5 and its version string ("1.10.6").
std::string library_description();

class File {
public:
    // mode "r": open read-only; "r+c": open read-only, creating an empty file
    // if none exists (HighFive's ReadOnly | Create); "w": create or truncate.
    File(const std::string& path, const std::string& mode);
    ~File();
    File(const File&) = delete;
    File& operator=(const File&) = delete;

    // Names of the datasets directly under the root group, in name order.
    std::vector<std::string> root_datasets() const;
    // Whole dataset converted to float64, flattened in storage order.
    std::vector<double> read(const std::string& name) const;

    // Creates a group under the root (fails if it exists).
    void create_group(const std::string& name);
    // Writes a float64 dataset of the given extents (a rank-1 dataset when
    // dims has one entry) at path "group/name" or "name".
    void write(const std::string& path, const std::vector<uint64_t>& dims, const double* data);

private:
    int64_t id_ = -1;
    std::string path_;
};

}  // namespace h5
}  // namespace ffip_amd
