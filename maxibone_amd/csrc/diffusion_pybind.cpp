// pybind11 module `diffusion`: the drop-in replacement for the reference's
// lib.cpp.{cpu,gpu}.diffusion module (src/pybind/diffusion-pybind.cc:81-86).
// A thin layer over the C ABI in include/maxibone_diffusion.h: it unpacks the
// numpy buffers, checks shapes, releases the GIL and maps status codes to
// Python exceptions.  It includes no HIP header and holds no device state.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <string>
#include <tuple>

#include "maxibone_diffusion.h"

namespace py = pybind11;

// Same buffer types as the reference (src/lib/cpp/include/datatypes.hh:56-57).
template <typename T>
using np_array = py::array_t<T, py::array::c_style | py::array::forcecast>;

namespace {

void raise_on(int rc) {
    if (rc == MXB_OK) return;
    const std::string msg = mxb_last_error();
    switch (rc) {
        case MXB_EINVAL: throw py::value_error(msg);
        case MXB_ENOMEM:
            PyErr_SetString(PyExc_MemoryError, msg.c_str());
            throw py::error_already_set();
        default: throw std::runtime_error(msg);
    }
}

struct Vol {
    int64_t nz, ny, nx;
};

Vol check_volume(const py::buffer_info &v) {
    if (v.ndim != 3) throw py::value_error("np_voxels must be 3-dimensional, got ndim=" + std::to_string(v.ndim));
    return {(int64_t)v.shape[0], (int64_t)v.shape[1], (int64_t)v.shape[2]};
}

int64_t check_kernel(const py::buffer_info &k) {
    if (k.ndim != 1) throw py::value_error("np_kernel must be 1-dimensional");
    if (k.shape[0] % 2 == 0) throw py::value_error("np_kernel must have an odd number of taps");
    return (int64_t)k.shape[0];
}

template <typename T>
void check_output(const py::buffer_info &o, const Vol &v, const char *name) {
    if (o.ndim != 3 || o.shape[0] != v.nz || o.shape[1] != v.ny || o.shape[2] != v.nx)
        throw py::value_error(std::string(name) + " must have the shape of np_voxels");
    if (o.readonly) throw py::value_error(std::string(name) + " must be writeable");
}

int precision_of(const std::string &p) {
    if (p == "fp32") return MXB_FP32;
    if (p == "bf16") return MXB_BF16;
    if (p == "fp16") return MXB_FP16;
    if (p == "fp32_fma") return MXB_FP32_FMA;
    if (p == "fp32_exact") return MXB_FP32_EXACT;
    if (p == "fp32_split") return MXB_FP32_SPLIT;
    throw py::value_error("precision must be one of fp32, bf16, fp16, fp32_fma, fp32_exact, fp32_split");
}

// Overload 1 (pyb:27-53, bound at pyb:84).
void diffusion_in_memory(np_array<uint8_t> &np_voxels, const np_array<float> &np_kernel,
                         np_array<uint16_t> &np_output, const int64_t repititions,
                         const int verbose) {
    auto vi = np_voxels.request(), ki = np_kernel.request(), oi = np_output.request(true);
    const Vol v = check_volume(vi);
    const int64_t K = check_kernel(ki);
    check_output<uint16_t>(oi, v, "np_output");
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = mxb_diffusion_u8_u16(static_cast<const uint8_t *>(vi.ptr), v.nz, v.ny, v.nx,
                                  static_cast<const float *>(ki.ptr), K, repititions,
                                  static_cast<uint16_t *>(oi.ptr), verbose);
    }
    raise_on(rc);
}

// Overload 2 (pyb:67-77, bound at pyb:85).
void diffusion_on_disk(const std::string &input_file, const np_array<float> &np_kernel,
                       const std::string &output_file,
                       const std::tuple<int64_t, int64_t, int64_t> &total_shape,
                       const std::tuple<int64_t, int64_t, int64_t> &global_shape,
                       const int64_t repititions, const int verbose) {
    auto ki = np_kernel.request();
    const int64_t K = check_kernel(ki);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = mxb_diffusion_on_disk(input_file.c_str(), static_cast<const float *>(ki.ptr), K,
                                   output_file.c_str(), std::get<0>(total_shape),
                                   std::get<1>(total_shape), std::get<2>(total_shape),
                                   std::get<0>(global_shape), std::get<1>(global_shape),
                                   std::get<2>(global_shape), repititions, verbose);
    }
    raise_on(rc);
}

// Extensions (no reference counterpart): FP32 output, precision forms, and
// the chunked form over host memory.
void diffusion_ex(np_array<uint8_t> &np_voxels, const np_array<float> &np_kernel,
                  py::object output_u16, py::object output_f32, const int64_t repititions,
                  const std::string &precision, const int verbose) {
    auto vi = np_voxels.request(), ki = np_kernel.request();
    const Vol v = check_volume(vi);
    const int64_t K = check_kernel(ki);
    const int prec = precision_of(precision);
    uint16_t *q = nullptr;
    float *f = nullptr;
    py::array_t<uint16_t, py::array::c_style> oq;
    py::array_t<float, py::array::c_style> of;
    if (!output_u16.is_none()) {
        oq = output_u16.cast<py::array_t<uint16_t, py::array::c_style>>();
        if (!py::isinstance<py::array_t<uint16_t, py::array::c_style>>(output_u16) || oq.ptr() != output_u16.ptr())
            throw py::type_error("output_u16 must be a C-contiguous uint16 array");
        auto oi = oq.request(true);
        check_output<uint16_t>(oi, v, "output_u16");
        q = static_cast<uint16_t *>(oi.ptr);
    }
    if (!output_f32.is_none()) {
        of = output_f32.cast<py::array_t<float, py::array::c_style>>();
        if (of.ptr() != output_f32.ptr()) throw py::type_error("output_f32 must be a C-contiguous float32 array");
        auto oi = of.request(true);
        check_output<float>(oi, v, "output_f32");
        f = static_cast<float *>(oi.ptr);
    }
    if (q == nullptr && f == nullptr) throw py::value_error("give output_u16 and/or output_f32");
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = mxb_diffusion_ex(static_cast<const uint8_t *>(vi.ptr), v.nz, v.ny, v.nx,
                              static_cast<const float *>(ki.ptr), K, repititions, prec, q, f, verbose);
    }
    raise_on(rc);
}

void diffusion_chunked(np_array<uint8_t> &np_voxels, const np_array<float> &np_kernel,
                       py::object output_u16, py::object output_f32, const int64_t repititions,
                       const int64_t chunk_z, const int64_t temporal, const std::string &precision,
                       const int verbose) {
    auto vi = np_voxels.request(), ki = np_kernel.request();
    const Vol v = check_volume(vi);
    const int64_t K = check_kernel(ki);
    const int prec = precision_of(precision);
    uint16_t *q = nullptr;
    float *f = nullptr;
    py::array_t<uint16_t, py::array::c_style> oq;
    py::array_t<float, py::array::c_style> of;
    if (!output_u16.is_none()) {
        oq = output_u16.cast<py::array_t<uint16_t, py::array::c_style>>();
        if (oq.ptr() != output_u16.ptr()) throw py::type_error("output_u16 must be a C-contiguous uint16 array");
        auto oi = oq.request(true);
        check_output<uint16_t>(oi, v, "output_u16");
        q = static_cast<uint16_t *>(oi.ptr);
    }
    if (!output_f32.is_none()) {
        of = output_f32.cast<py::array_t<float, py::array::c_style>>();
        if (of.ptr() != output_f32.ptr()) throw py::type_error("output_f32 must be a C-contiguous float32 array");
        auto oi = of.request(true);
        check_output<float>(oi, v, "output_f32");
        f = static_cast<float *>(oi.ptr);
    }
    if (q == nullptr && f == nullptr) throw py::value_error("give output_u16 and/or output_f32");
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = mxb_diffusion_chunked(static_cast<const uint8_t *>(vi.ptr), v.nz, v.ny, v.nx,
                                   static_cast<const float *>(ki.ptr), K, repititions, chunk_z,
                                   temporal, prec, q, f, verbose);
    }
    raise_on(rc);
}

}  // namespace

PYBIND11_MODULE(diffusion, m) {
    m.doc() = "Diffusion approximation using 3 1D gauss convolutions (MI355X / gfx950).";
    m.def("diffusion", &diffusion_in_memory, py::arg("np_voxels"), py::arg("np_kernel"),
          py::arg("np_output").noconvert(), py::arg("repititions") = 1, py::arg("verbose") = 0);
    m.def("diffusion", &diffusion_on_disk, py::arg("input_file"), py::arg("np_kernel"),
          py::arg("output_file"), py::arg("total_shape"), py::arg("global_shape"),
          py::arg("repititions") = 1, py::arg("verbose") = 0);
    m.def("diffusion_ex", &diffusion_ex, py::arg("np_voxels"), py::arg("np_kernel"),
          py::arg("output_u16") = py::none(), py::arg("output_f32") = py::none(),
          py::arg("repititions") = 1, py::arg("precision") = "fp32", py::arg("verbose") = 0);
    m.def("diffusion_chunked", &diffusion_chunked, py::arg("np_voxels"), py::arg("np_kernel"),
          py::arg("output_u16") = py::none(), py::arg("output_f32") = py::none(),
          py::arg("repititions") = 1, py::arg("chunk_z") = 256, py::arg("temporal") = 1,
          py::arg("precision") = "fp32", py::arg("verbose") = 0);
    m.def("version", &mxb_version);
}
