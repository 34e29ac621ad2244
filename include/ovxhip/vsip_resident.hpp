// vsip_resident.hpp -- a VSIPL++ block whose data live in host memory and
// in HBM, for the gfx950 backend's evaluators.
//
// The reference's storage_manager keeps a block's data in host memory plus,
// when configured with OpenCL, a device copy, each valid or stale
// (src/ovxx/storage/manager.hpp:247-350; src/ovxx/opencl/storage.hpp:47-130),
// and its OpenCL DDA hands kernels the device buffer of blocks that have one
// (src/ovxx/opencl/dda.hpp:54-80).  A HIP location inside that manager is a
// change to the OpenVSIP tree; this header gets the same effect from outside
// it: `Resident<D, T>` is a dense row-major block (D = 1 or 2, T =
// complex<float> or float) whose storage is an ovxhip::Storage (the same
// valid/stale rules, with an HBM location).  Views over it behave as over a
// Dense block -- element access and DDA see host memory, kept up to date --
// while the plugin's evaluators ask it for device pointers instead, so an
// expression such as  out = inv_fftm(vmmul<0>(w, fwd_fftm(in)))  over
// Resident blocks runs on HBM-resident data with no host copy.
//
//   typedef ovxx::ovxhip_backend::Resident<2> rblock;
//   vsip::Matrix<vsip::complex<float>, rblock> in(rows, n), out(rows, n);
//
// Included by vsip_plugin.hpp (a configured OpenVSIP tree is needed).
#ifndef OVXHIP_VSIP_RESIDENT_HPP
#define OVXHIP_VSIP_RESIDENT_HPP

#include <ovxx/block_traits.hpp>
#include <ovxx/refcounted.hpp>
#include <vsip/domain.hpp>
#include <vsip/impl/local_map.hpp>
#include <vsip/layout.hpp>

#include "ovxhip.hpp"

namespace ovxx
{
namespace ovxhip_backend
{

template <dimension_type D>
struct row_major;
template <>
struct row_major<1>
{
  typedef vsip::row1_type type;
};
template <>
struct row_major<2>
{
  typedef vsip::row2_type type;
};

template <dimension_type D, typename T = vsip::complex<float> >
class Resident : public refcounted<Resident<D, T> >
{
  static_assert(D == 1 || D == 2, "Resident blocks are vectors or matrices");

public:
  typedef T value_type;
  typedef T &reference_type;
  typedef T const &const_reference_type;
  typedef vsip::Local_map map_type;
  typedef typename row_major<D>::type order_type;
  typedef vsip::Layout<D, order_type, vsip::dense, vsip::array> layout_type;
  typedef T *ptr_type;
  typedef T const *const_ptr_type;
  static dimension_type const dim = D;
  static vsip::storage_format_type const storage_format = vsip::array;

  explicit Resident(vsip::Domain<D> const &dom, map_type const & = map_type())
    : rows_(D == 2 ? dom[0].size() : 1), cols_(D == 2 ? dom[D - 1].size() : dom[0].size()),
      storage_((long)(rows_ * cols_))
  {}
  Resident(vsip::Domain<D> const &dom, T value, map_type const &map = map_type()) : Resident(dom, map)
  {
    T *p = storage_.host(OVX_ACCESS_WRITE);
    for (length_type i = 0; i < rows_ * cols_; ++i) p[i] = value;
  }

  length_type size() const { return rows_ * cols_; }
  length_type size(dimension_type block_dim, dimension_type d) const
  {
    if (block_dim == 1) return size();
    return d == 0 ? rows_ : cols_;
  }
  // dense row-major strides, in elements
  stride_type stride(dimension_type block_dim, dimension_type d) const
  {
    if (block_dim == 1 || D == 1) return 1;
    return d == 0 ? (stride_type)cols_ : 1;
  }

  // element access through the host location (refreshed when stale; a
  // write makes the device copy stale), as storage_manager::get/put
  T get(index_type i) const { return cstorage().ptr(OVX_LOC_HOST)[i]; }
  T get(index_type i, index_type j) const { return cstorage().ptr(OVX_LOC_HOST)[i * cols_ + j]; }
  void put(index_type i, T v) { storage_.host(OVX_ACCESS_READ_WRITE)[i] = v; }
  void put(index_type i, index_type j, T v) { storage_.host(OVX_ACCESS_READ_WRITE)[i * cols_ + j] = v; }

  // DDA: host pointers (storage_manager::ptr, location 0)
  ptr_type ptr() { return storage_.host(OVX_ACCESS_READ_WRITE); }
  const_ptr_type ptr() const { return cstorage().ptr(OVX_LOC_HOST); }

  // The backend's view: device pointers in HBM.  `access` is
  // OVX_ACCESS_READ, OVX_ACCESS_WRITE (no copy in, host goes stale) or
  // OVX_ACCESS_READ_WRITE.  A stale device copy is refreshed on `stream`.
  T *device(int access, void *stream = nullptr) { return storage_.device(access, stream); }
  T const *device(void *stream = nullptr) const { return cstorage().ptr(OVX_LOC_DEVICE, stream); }
  unsigned valid() const { return cstorage().valid(); }

  map_type const &map() const
  {
    static map_type const m;
    return m;
  }
  // no user storage: admit / release are the identity (manager.hpp:207-243)
  void admit(bool = true) {}
  void release(bool = true) {}
  bool admitted() const { return true; }
  vsip::user_storage_type user_storage() const { return vsip::no_user_format; }

private:
  // read-only requests go through the const overloads (no location is
  // made stale)
  ovxhip::Storage<T> const &cstorage() const { return storage_; }

  length_type rows_, cols_;
  ovxhip::Storage<T> storage_;
};

template <typename B>
struct is_resident_block
{
  static bool const value = false;
};
template <typename B>
struct is_resident_block<B const> : is_resident_block<B>
{};
template <dimension_type D, typename T>
struct is_resident_block<Resident<D, T> >
{
  static bool const value = true;
};

template <dimension_type D, typename T>
Resident<D, T> &get_local_block(Resident<D, T> &block)
{
  return block;
}
template <dimension_type D, typename T>
Resident<D, T> const &get_local_block(Resident<D, T> const &block)
{
  return block;
}
template <dimension_type D, typename T>
void assert_local(Resident<D, T> const &, index_type sb)
{
  OVXX_PRECONDITION(sb == 0);
}

} // namespace ovxx::ovxhip_backend

template <dimension_type D, typename T>
struct is_modifiable_block<ovxhip_backend::Resident<D, T> >
{
  static bool const value = true;
};

template <dimension_type D, typename T>
struct distributed_local_block<ovxhip_backend::Resident<D, T> >
{
  typedef ovxhip_backend::Resident<D, T> type;
  typedef ovxhip_backend::Resident<D, T> proxy_type;
};

} // namespace ovxx

namespace vsip
{
template <dimension_type D, typename T>
struct get_block_layout<ovxx::ovxhip_backend::Resident<D, T> >
{
  static dimension_type const dim = D;
  typedef typename ovxx::ovxhip_backend::Resident<D, T>::order_type order_type;
  static pack_type const packing = dense;
  static storage_format_type const storage_format = array;
  typedef Layout<D, order_type, packing, storage_format> type;
};

template <dimension_type D, typename T>
struct supports_dda<ovxx::ovxhip_backend::Resident<D, T> >
{
  static bool const value = true;
};
} // namespace vsip

#endif
