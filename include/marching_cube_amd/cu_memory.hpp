/*
 * cu_memory.hpp -- cu_array<T>: a reference-counted HIP device buffer, or a
 * non-owning view of one (the part of the reference's cu_memory.hpp:373-460,
 * 720 that the mesh path uses: create, release, empty, size, ptr, upload,
 * download, and the non-owning cu_array(T*, n) constructor).
 */
#ifndef MARCHING_CUBE_AMD_CU_MEMORY_HPP
#define MARCHING_CUBE_AMD_CU_MEMORY_HPP

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>

namespace mcx_detail {
inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace mcx_detail

template <typename T>
class cu_array {
 public:
  cu_array() = default;
  explicit cu_array(size_t n) { create(n); }
  // non-owning view of n elements at a device pointer
  cu_array(T* ptr, size_t n) : ptr_(ptr), size_(n) {}

  void create(size_t n) {
    release();
    if (n == 0) return;
    void* p = nullptr;
    mcx_detail::check_hip(hipMalloc(&p, n * sizeof(T)), "cu_array::create");
    owner_ = std::shared_ptr<void>(p, [](void* q) { (void)hipFree(q); });
    ptr_ = static_cast<T*>(p);
    size_ = n;
  }
  void release() {
    owner_.reset();
    ptr_ = nullptr;
    size_ = 0;
  }
  bool empty() const { return ptr_ == nullptr || size_ == 0; }
  size_t size() const { return size_; }
  size_t sizeBytes() const { return size_ * sizeof(T); }
  T* ptr() { return ptr_; }
  const T* ptr() const { return ptr_; }
  operator T*() { return ptr_; }
  operator const T*() const { return ptr_; }

  void upload(const T* host, size_t n) {
    if (size_ != n) create(n);
    if (n)
      mcx_detail::check_hip(hipMemcpy(ptr_, host, n * sizeof(T), hipMemcpyHostToDevice),
                            "cu_array::upload");
  }
  void download(T* host) const {
    if (size_)
      mcx_detail::check_hip(hipMemcpy(host, ptr_, size_ * sizeof(T), hipMemcpyDeviceToHost),
                            "cu_array::download");
  }

 private:
  std::shared_ptr<void> owner_;
  T* ptr_ = nullptr;
  size_t size_ = 0;
};

#endif
