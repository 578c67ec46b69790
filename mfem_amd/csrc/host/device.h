// Device memory and stream plumbing for the host framework.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace mfem {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define MFEM_HIP_CHECK(expr)                                                              \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      throw ::mfem::HipError(std::string(#expr) + ": " + hipGetErrorString(_e) + " (" + \
                             __FILE__ + ":" + std::to_string(__LINE__) + ")");            \
  } while (0)

// Owning device array (hipMalloc / hipFree).
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  ~DeviceArray() { release(); }

  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n > 0) MFEM_HIP_CHECK(hipMalloc(&p_, n * sizeof(T)));
    n_ = n;
  }
  void upload(const T* h, size_t n, hipStream_t st = nullptr) {
    resize(n);
    if (n) MFEM_HIP_CHECK(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, st));
  }
  void upload(const std::vector<T>& h, hipStream_t st = nullptr) { upload(h.data(), h.size(), st); }
  void download(T* h, hipStream_t st = nullptr) const {
    if (n_) {
      MFEM_HIP_CHECK(hipMemcpyAsync(h, p_, n_ * sizeof(T), hipMemcpyDeviceToHost, st));
      MFEM_HIP_CHECK(hipStreamSynchronize(st));
    }
  }
  std::vector<T> to_host(hipStream_t st = nullptr) const {
    std::vector<T> v(n_);
    download(v.data(), st);
    return v;
  }
  void zero(hipStream_t st = nullptr) {
    if (n_) MFEM_HIP_CHECK(hipMemsetAsync(p_, 0, n_ * sizeof(T), st));
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Precision-erased element array: FP64 or FP32 storage chosen at runtime.
class RealArray {
 public:
  void resize(size_t n, bool fp32) {
    fp32_ = fp32;
    if (fp32) { f_.resize(n); d_.resize(0); } else { d_.resize(n); f_.resize(0); }
  }
  bool fp32() const { return fp32_; }
  template <typename T> T* as();
  template <typename T> const T* as() const;
  float* f32() { return f_.data(); }
  double* f64() { return d_.data(); }
  const float* f32() const { return f_.data(); }
  const double* f64() const { return d_.data(); }
  size_t bytes() const { return fp32_ ? f_.bytes() : d_.bytes(); }
  std::vector<double> to_host_double(hipStream_t st = nullptr) const {
    if (!fp32_) return d_.to_host(st);
    std::vector<float> f = f_.to_host(st);
    return std::vector<double>(f.begin(), f.end());
  }

 private:
  bool fp32_ = false;
  DeviceArray<double> d_;
  DeviceArray<float> f_;
};
template <> inline double* RealArray::as<double>() { return d_.data(); }
template <> inline float* RealArray::as<float>() { return f_.data(); }
template <> inline const double* RealArray::as<double>() const { return d_.data(); }
template <> inline const float* RealArray::as<float>() const { return f_.data(); }

// Small pinned host buffer for scalar read-backs.
class PinnedBuffer {
 public:
  // flags: hipHostMalloc flags; with hipHostMallocMapped, device() is the
  // address kernels write the buffer through
  explicit PinnedBuffer(size_t bytes = 256, unsigned flags = hipHostMallocDefault) {
    MFEM_HIP_CHECK(hipHostMalloc(&p_, bytes, flags));
    if (flags & hipHostMallocMapped) MFEM_HIP_CHECK(hipHostGetDevicePointer(&d_, p_, 0));
  }
  ~PinnedBuffer() { if (p_) (void)hipHostFree(p_); }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  template <typename T> T* as() { return reinterpret_cast<T*>(p_); }
  template <typename T> T* device() { return reinterpret_cast<T*>(d_); }

 private:
  void* p_ = nullptr;
  void* d_ = nullptr;
};

}  // namespace mfem
