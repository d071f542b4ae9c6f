// HDF5 access through libhdf5's public C API, resolved with dlopen/dlsym.
// The declarations below are the documented HDF5 1.10+ signatures (hid_t is
// 64-bit from 1.10 on); only the calls this file makes are declared.
#include "hdf5_io.hpp"

#include <dlfcn.h>
#include <sys/stat.h>

#include <cstdlib>
#include <mutex>
#include <stdexcept>

namespace ffip_amd {
namespace h5 {
namespace {

using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
using hssize_t = long long;

constexpr hid_t P_DEFAULT = 0;  // H5P_DEFAULT
constexpr hid_t S_ALL = 0;      // H5S_ALL
constexpr hid_t E_DEFAULT = 0;  // H5E_DEFAULT
constexpr unsigned ACC_RDONLY = 0x0u, ACC_TRUNC = 0x2u, ACC_EXCL = 0x4u;
constexpr int INDEX_NAME = 0, ITER_INC = 0;  // H5_INDEX_NAME, H5_ITER_INC

using link_cb = herr_t (*)(hid_t, const char*, const void*, void*);

struct Api {
    void* handle = nullptr;
    std::string path;
    unsigned ver[3] = {0, 0, 0};
    herr_t (*open)();
    herr_t (*get_libversion)(unsigned*, unsigned*, unsigned*);
    herr_t (*eset_auto)(hid_t, void*, void*);
    hid_t (*fopen)(const char*, unsigned, hid_t);
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gcreate)(hid_t, const char*, hid_t, hid_t, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t (*dopen)(hid_t, const char*, hid_t);
    hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dget_space)(hid_t);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*dclose)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
    hssize_t (*sget_npoints)(hid_t);
    herr_t (*sclose)(hid_t);
    herr_t (*literate)(hid_t, int, int, hsize_t*, link_cb, void*);
    hid_t* native_double;  // H5T_NATIVE_DOUBLE_g
    hid_t* ieee_f64le;     // H5T_IEEE_F64LE_g
};

template <class F>
void bind(Api& a, F& fn, const char* name)
{
    void* s = dlsym(a.handle, name);
    if (!s)
        throw std::runtime_error(a.path + " does not export " + name);
    fn = reinterpret_cast<F>(s);
}

const Api& api()
{
    static Api a;
    static std::string err;
    static std::once_flag once;
    std::call_once(once, [] {
        std::vector<std::string> cands;
        if (const char* e = std::getenv("FFIP_HDF5_LIBRARY"))
            cands.push_back(e);
        cands.insert(cands.end(), {"libhdf5.so", "libhdf5_serial.so", "/opt/conda/lib/libhdf5.so"});
        for (auto& c : cands) {
            a.handle = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (a.handle) {
                a.path = c;
                break;
            }
        }
        if (!a.handle) {
            err = "libhdf5 was not found (set FFIP_HDF5_LIBRARY to its path)";
            return;
        }
        try {
            bind(a, a.open, "H5open");
            bind(a, a.get_libversion, "H5get_libversion");
            bind(a, a.eset_auto, "H5Eset_auto2");
            bind(a, a.fopen, "H5Fopen");
            bind(a, a.fcreate, "H5Fcreate");
            bind(a, a.fclose, "H5Fclose");
            bind(a, a.gcreate, "H5Gcreate2");
            bind(a, a.gclose, "H5Gclose");
            bind(a, a.dopen, "H5Dopen2");
            bind(a, a.dcreate, "H5Dcreate2");
            bind(a, a.dget_space, "H5Dget_space");
            bind(a, a.dread, "H5Dread");
            bind(a, a.dwrite, "H5Dwrite");
            bind(a, a.dclose, "H5Dclose");
            bind(a, a.screate_simple, "H5Screate_simple");
            bind(a, a.sget_npoints, "H5Sget_simple_extent_npoints");
            bind(a, a.sclose, "H5Sclose");
            // 1.12 renamed the iterator; both forms take the same arguments
            if (dlsym(a.handle, "H5Literate2"))
                bind(a, a.literate, "H5Literate2");
            else
                bind(a, a.literate, "H5Literate");
            bind(a, a.native_double, "H5T_NATIVE_DOUBLE_g");
            bind(a, a.ieee_f64le, "H5T_IEEE_F64LE_g");
            if (a.open() < 0)
                throw std::runtime_error("H5open failed");
            a.get_libversion(&a.ver[0], &a.ver[1], &a.ver[2]);
            if (a.ver[0] < 1 || (a.ver[0] == 1 && a.ver[1] < 10))
                throw std::runtime_error(a.path + " is older than HDF5 1.10 (32-bit identifiers)");
            a.eset_auto(E_DEFAULT, nullptr, nullptr);  // errors are reported by exceptions
        } catch (std::exception& e) {
            err = e.what();
        }
    });
    if (!err.empty())
        throw std::runtime_error(err);
    return a;
}

bool exists(const std::string& p)
{
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

}  // namespace

std::string library_description()
{
    const Api& a = api();
    return a.path + " (HDF5 " + std::to_string(a.ver[0]) + "." + std::to_string(a.ver[1]) + "." +
           std::to_string(a.ver[2]) + ")";
}

File::File(const std::string& path, const std::string& mode) : path_(path)
{
    const Api& a = api();
    if (mode == "r" || (mode == "r+c" && exists(path))) {
        id_ = a.fopen(path.c_str(), ACC_RDONLY, P_DEFAULT);
    } else if (mode == "r+c") {
        id_ = a.fcreate(path.c_str(), ACC_EXCL, P_DEFAULT, P_DEFAULT);
    } else if (mode == "w") {
        id_ = a.fcreate(path.c_str(), ACC_TRUNC, P_DEFAULT, P_DEFAULT);
    } else {
        throw std::runtime_error("HDF5 file mode must be r, r+c or w");
    }
    if (id_ < 0)
        throw std::runtime_error("cannot open HDF5 file \"" + path + "\"");
}

File::~File()
{
    if (id_ >= 0)
        api().fclose(id_);
}

std::vector<std::string> File::root_datasets() const
{
    const Api& a = api();
    std::vector<std::string> names;
    hsize_t idx = 0;
    link_cb cb = [](hid_t, const char* name, const void*, void* data) -> herr_t {
        static_cast<std::vector<std::string>*>(data)->push_back(name);
        return 0;
    };
    if (a.literate(id_, INDEX_NAME, ITER_INC, &idx, cb, &names) < 0)
        throw std::runtime_error("cannot list the root group of \"" + path_ + "\"");
    std::vector<std::string> out;
    for (auto& n : names) {
        hid_t d = a.dopen(id_, n.c_str(), P_DEFAULT);
        if (d >= 0) {  // groups and other objects fail to open as datasets
            out.push_back(n);
            a.dclose(d);
        }
    }
    return out;
}

std::vector<double> File::read(const std::string& name) const
{
    const Api& a = api();
    hid_t d = a.dopen(id_, name.c_str(), P_DEFAULT);
    if (d < 0)
        throw std::runtime_error("dataset \"" + name + "\" not found in \"" + path_ + "\"");
    hid_t s = a.dget_space(d);
    hssize_t n = s >= 0 ? a.sget_npoints(s) : -1;
    if (s >= 0)
        a.sclose(s);
    std::vector<double> v(n > 0 ? size_t(n) : 0);
    herr_t rc = n < 0 ? -1 : n == 0 ? 0 : a.dread(d, *a.native_double, S_ALL, S_ALL, P_DEFAULT, v.data());
    a.dclose(d);
    if (rc < 0)
        throw std::runtime_error("cannot read dataset \"" + name + "\" as float64");
    return v;
}

void File::create_group(const std::string& name)
{
    const Api& a = api();
    hid_t g = a.gcreate(id_, name.c_str(), P_DEFAULT, P_DEFAULT, P_DEFAULT);
    if (g < 0)
        throw std::runtime_error("cannot create group \"" + name + "\" in \"" + path_ + "\"");
    a.gclose(g);
}

void File::write(const std::string& path, const std::vector<uint64_t>& dims, const double* data)
{
    const Api& a = api();
    std::vector<hsize_t> hd(dims.begin(), dims.end());
    hsize_t total = 1;
    for (auto x : hd)
        total *= x;
    hid_t s = a.screate_simple(int(hd.size()), hd.data(), nullptr);
    if (s < 0)
        throw std::runtime_error("cannot create a dataspace for \"" + path + "\"");
    hid_t d = a.dcreate(id_, path.c_str(), *a.ieee_f64le, s, P_DEFAULT, P_DEFAULT, P_DEFAULT);
    herr_t rc = d < 0 ? -1 : total == 0 ? 0 : a.dwrite(d, *a.native_double, S_ALL, S_ALL, P_DEFAULT, data);
    if (d >= 0)
        a.dclose(d);
    a.sclose(s);
    if (rc < 0)
        throw std::runtime_error("cannot write dataset \"" + path + "\" in \"" + path_ + "\"");
}

}  // namespace h5
}  // namespace ffip_amd
