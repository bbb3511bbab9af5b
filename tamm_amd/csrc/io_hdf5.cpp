// io_hdf5.cpp -- tensor I/O in TAMM's on-disk format (write_to_disk / read_from_disk,
// src/tamm/tamm_io.hpp:223-460 and :735-980, the tammio path).
//
// Format: one HDF5 file holding a 1-D dataset "tensor" of native doubles whose extent is the
// tensor's dense element count (product of the modes' index counts).  Nonzero blocks are stored
// back to back from offset 0, each at the sum of the sizes of the nonzero blocks before it in
// row-major block-id order -- which is exactly the Distribution_NW buffer layout
// (distribution.hpp:197-365) the engine keeps in HBM, so a tensor goes to disk with a single
// hyperslab write of its packed buffer and comes back with a single read.
//
// libhdf5 is loaded with dlopen on first use (HDF5 1.10 ABI: hid_t is int64_t), so the engine
// has no link-time dependency on it; without the library these entry points return
// TMX_ERR_UNSUPPORTED and everything else is unaffected.
#include <dlfcn.h>

#include <mutex>

#include "tmx_internal.hpp"

namespace tmx {
namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr hid_t kH5PDefault = 0;
constexpr unsigned kH5FAccRdonly = 0x0000u, kH5FAccTrunc = 0x0002u;
constexpr int kH5SSelectSet = 0;

struct H5Api {
  bool ok = false;
  herr_t (*open)();
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char*, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
  herr_t (*sclose)(hid_t);
  herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t*, const hsize_t*, const hsize_t*, const hsize_t*);
  int (*sget_simple_extent_dims)(hid_t, hsize_t*, hsize_t*);
  int (*sget_simple_extent_ndims)(hid_t);
  hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen2)(hid_t, const char*, hid_t);
  herr_t (*dclose)(hid_t);
  hid_t (*dget_space)(hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
  hid_t* native_double;
};

template <typename F>
bool sym(void* h, const char* name, F& f) {
  f = reinterpret_cast<F>(dlsym(h, name));
  return f != nullptr;
}

const H5Api& h5() {
  static H5Api api;
  static std::once_flag once;
  std::call_once(once, [] {
    const char* names[] = {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
    void* h = nullptr;
    for (const char* n : names)
      if ((h = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) return;
    bool ok = sym(h, "H5open", api.open) && sym(h, "H5Fcreate", api.fcreate) &&
              sym(h, "H5Fopen", api.fopen) && sym(h, "H5Fclose", api.fclose) &&
              sym(h, "H5Screate_simple", api.screate_simple) && sym(h, "H5Sclose", api.sclose) &&
              sym(h, "H5Sselect_hyperslab", api.sselect_hyperslab) &&
              sym(h, "H5Sget_simple_extent_dims", api.sget_simple_extent_dims) &&
              sym(h, "H5Sget_simple_extent_ndims", api.sget_simple_extent_ndims) &&
              sym(h, "H5Dcreate2", api.dcreate2) && sym(h, "H5Dopen2", api.dopen2) &&
              sym(h, "H5Dclose", api.dclose) && sym(h, "H5Dget_space", api.dget_space) &&
              sym(h, "H5Dwrite", api.dwrite) && sym(h, "H5Dread", api.dread) &&
              sym(h, "H5T_NATIVE_DOUBLE_g", api.native_double);
    if (ok && api.open() >= 0) api.ok = true;
  });
  if (!api.ok) throw Error(TMX_ERR_UNSUPPORTED, "HDF5 (libhdf5.so, 1.10 ABI) could not be loaded");
  return api;
}

// Closes HDF5 handles in reverse order of acquisition on every exit path.
struct H5Handles {
  const H5Api& a;
  hid_t file = -1, space = -1, dset = -1, fspace = -1, mspace = -1;
  ~H5Handles() {
    if (mspace >= 0) a.sclose(mspace);
    if (fspace >= 0) a.sclose(fspace);
    if (dset >= 0) a.dclose(dset);
    if (space >= 0) a.sclose(space);
    if (file >= 0) a.fclose(file);
  }
};

// tmx_tensor handles point at the engine's shared_ptr<Tensor> (structure.cpp).
const Tensor& io_tensor(tmx_tensor t) {
  TMX_REQUIRE(t, "null tensor");
  return **reinterpret_cast<std::shared_ptr<Tensor>*>(t);
}

hsize_t dense_count(const Tensor& t) {
  hsize_t n = 1;
  for (const auto& m : t.modes) n *= (hsize_t)m->n;
  return n;
}

}  // namespace
}  // namespace tmx

using namespace tmx;

extern "C" {

int tmx_tensor_write_hdf5(tmx_tensor t, const double* data, const char* path, void* stream) {
  return guarded([&] {
    require_init();
    TMX_REQUIRE(t && data && path, "null argument");
    const Tensor& tt = io_tensor(t);
    const H5Api& a = h5();
    std::vector<double> host((size_t)tt.total);
    hipStream_t s = pick_stream(stream);
    if (tt.total > 0)
      TMX_HIP(hipMemcpyAsync(host.data(), data, sizeof(double) * tt.total, hipMemcpyDeviceToHost, s));
    TMX_HIP(hipStreamSynchronize(s));
    H5Handles h{a};
    h.file = a.fcreate(path, kH5FAccTrunc, kH5PDefault, kH5PDefault);
    TMX_REQUIRE(h.file >= 0, std::string("cannot create ") + path);
    const hsize_t dims = dense_count(tt);
    h.space = a.screate_simple(1, &dims, nullptr);
    h.dset = a.dcreate2(h.file, "tensor", *a.native_double, h.space, kH5PDefault, kH5PDefault, kH5PDefault);
    TMX_REQUIRE(h.space >= 0 && h.dset >= 0, "cannot create dataset 'tensor'");
    if (tt.total > 0) {
      const hsize_t off = 0, stride = 1, cnt = (hsize_t)tt.total;
      h.fspace = a.dget_space(h.dset);
      TMX_REQUIRE(a.sselect_hyperslab(h.fspace, kH5SSelectSet, &off, &stride, &cnt, nullptr) >= 0,
                  "hyperslab selection failed");
      h.mspace = a.screate_simple(1, &cnt, nullptr);
      TMX_REQUIRE(a.dwrite(h.dset, *a.native_double, h.mspace, h.fspace, kH5PDefault, host.data()) >= 0,
                  "H5Dwrite failed");
    }
  });
}

int tmx_tensor_read_hdf5(tmx_tensor t, double* data, const char* path, void* stream) {
  return guarded([&] {
    require_init();
    TMX_REQUIRE(t && data && path, "null argument");
    const Tensor& tt = io_tensor(t);
    const H5Api& a = h5();
    H5Handles h{a};
    h.file = a.fopen(path, kH5FAccRdonly, kH5PDefault);
    TMX_REQUIRE(h.file >= 0, std::string("cannot open ") + path);
    h.dset = a.dopen2(h.file, "tensor", kH5PDefault);
    TMX_REQUIRE(h.dset >= 0, "file has no dataset 'tensor'");
    h.fspace = a.dget_space(h.dset);
    hsize_t dims = 0;
    TMX_REQUIRE(a.sget_simple_extent_ndims(h.fspace) == 1 &&
                    a.sget_simple_extent_dims(h.fspace, &dims, nullptr) == 1,
                "dataset 'tensor' is not one-dimensional");
    TMX_REQUIRE(dims == dense_count(tt),
                "dataset 'tensor' does not match the tensor's dense size (different index spaces)");
    std::vector<double> host((size_t)tt.total);
    if (tt.total > 0) {
      const hsize_t off = 0, stride = 1, cnt = (hsize_t)tt.total;
      TMX_REQUIRE(a.sselect_hyperslab(h.fspace, kH5SSelectSet, &off, &stride, &cnt, nullptr) >= 0,
                  "hyperslab selection failed");
      h.mspace = a.screate_simple(1, &cnt, nullptr);
      TMX_REQUIRE(a.dread(h.dset, *a.native_double, h.mspace, h.fspace, kH5PDefault, host.data()) >= 0,
                  "H5Dread failed");
      hipStream_t s = pick_stream(stream);
      TMX_HIP(hipMemcpyAsync(data, host.data(), sizeof(double) * tt.total, hipMemcpyHostToDevice, s));
      TMX_HIP(hipStreamSynchronize(s));
    }
  });
}

}  // extern "C"
