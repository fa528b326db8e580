// C++ layer of the MI355X NCC find path: mvkit's C++ API for NCC matching
// (include/mvkit/cpp/mv_ncc_matching.hpp:10-482 and the support classes it
// names: mv_status.hpp, mv_api.hpp MvException, mv_string.hpp, mv_tuple.hpp,
// mv_handle.hpp, mv_iconic.hpp, mv_image.hpp, mv_region.hpp), header-only over
// the C ABI of mv_ncc.h.
//
// Semantics kept from the reference's C++ layer (src/cpp_api/ncc_matching.cpp,
// src/cpp_api/api_bridge.hpp):
//  - every failing C call throws mv::MvException carrying the C ABI's last
//    error (status, code, operation, parameter, parameter index, message);
//  - outputs are committed together after the call succeeded; a null or
//    repeated output pointer is an InvalidArgument before any work;
//  - copies of a handle (MvNCCModel, MvSerializedItem, MvHandle) share one
//    model; ClearHandle on any copy invalidates all of them;
//  - iconic objects (MvImage, MvRegion) are immutable once built and shared
//    by copies.
#ifndef MVKIT_AMD_MV_NCC_HPP
#define MVKIT_AMD_MV_NCC_HPP

#include <cstdint>
#include <cstring>
#include <exception>
#include <initializer_list>
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "mv_ncc.h"

namespace mv {

// ------------------------------------------------------------- status codes
// include/mvkit/cpp/mv_status.hpp
enum class MvStatus : std::int32_t {
  Ok = MV_STATUS_OK,
  InvalidArgument = MV_STATUS_INVALID_ARGUMENT,
  OutOfRange = MV_STATUS_OUT_OF_RANGE,
  TypeMismatch = MV_STATUS_TYPE_MISMATCH,
  NotFound = MV_STATUS_NOT_FOUND,
  OutOfMemory = MV_STATUS_OUT_OF_MEMORY,
  Unsupported = MV_STATUS_UNSUPPORTED,
  InvalidState = MV_STATUS_INVALID_STATE,
  IoError = MV_STATUS_IO_ERROR,
  DecodeError = MV_STATUS_DECODE_ERROR,
  EncodeError = MV_STATUS_ENCODE_ERROR,
  LimitExceeded = MV_STATUS_LIMIT_EXCEEDED,
  InternalError = MV_STATUS_INTERNAL_ERROR,
};
constexpr bool succeeded(MvStatus status) noexcept { return static_cast<std::int32_t>(status) >= 0; }
constexpr bool failed(MvStatus status) noexcept { return !succeeded(status); }

// include/mvkit/c/mv_api.h:57-72
enum class MvErrorCode : std::int32_t {
  Unspecified = MV_ERROR_UNSPECIFIED,
  NullArgument = MV_ERROR_NULL_ARGUMENT,
  InvalidCardinality = MV_ERROR_INVALID_CARDINALITY,
  InvalidType = MV_ERROR_INVALID_TYPE,
  InvalidValue = MV_ERROR_INVALID_VALUE,
  IndexOutOfRange = MV_ERROR_INDEX_OUT_OF_RANGE,
  MissingItem = MV_ERROR_MISSING_ITEM,
  UnsupportedValue = MV_ERROR_UNSUPPORTED_VALUE,
  InvalidObjectState = MV_ERROR_INVALID_OBJECT_STATE,
  IoFailure = MV_ERROR_IO_FAILURE,
  InternalFailure = MV_ERROR_INTERNAL_FAILURE,
};

// include/mvkit/c/mv_handle.h:18-29 (the kinds this library creates)
enum class MvObjectType : std::uint64_t { Unknown = 0, NccModel = 9, SerializedItem = 12 };

// ---------------------------------------------------------------- exception
// include/mvkit/cpp/mv_api.hpp MvException.  Copies share the immutable text.
class MvException final : public std::exception {
 public:
  MvException(MvStatus status, MvErrorCode code, ::MvStringView operation, ::MvStringView parameter,
              std::int64_t index, ::MvStringView message) {
    auto d = std::make_shared<Data>();
    d->status = status;
    d->code = code;
    d->operation = text(operation);
    d->parameter = text(parameter);
    d->index = index;
    d->message = text(message);
    d->what = d->operation;
    if (!d->parameter.empty()) {
      d->what += " [" + d->parameter;
      if (index >= 0) d->what += "[" + std::to_string(index) + "]";
      d->what += "]";
    }
    d->what += ": " + d->message + " (status " + std::to_string(static_cast<std::int32_t>(status)) + ")";
    d_ = std::move(d);
  }
  MvException(MvStatus status, MvErrorCode code, const char* operation, const char* parameter, std::int64_t index,
              const char* message)
      : MvException(status, code, view(operation), view(parameter), index, view(message)) {}

  MvStatus GetStatus() const noexcept { return d_->status; }
  MvErrorCode GetErrorCode() const noexcept { return d_->code; }
  ::MvStringView GetOperatorName() const noexcept { return {d_->operation.data(), d_->operation.size()}; }
  ::MvStringView GetParameterName() const noexcept { return {d_->parameter.data(), d_->parameter.size()}; }
  std::int64_t GetParameterIndex() const noexcept { return d_->index; }
  ::MvStringView GetMessage() const noexcept { return {d_->message.data(), d_->message.size()}; }
  const char* what() const noexcept override { return d_->what.c_str(); }

 private:
  static ::MvStringView view(const char* s) noexcept { return {s ? s : "", s ? std::strlen(s) : 0}; }
  static std::string text(::MvStringView v) { return v.data ? std::string(v.data, v.size) : std::string(); }
  struct Data {
    MvStatus status = MvStatus::Ok;
    MvErrorCode code = MvErrorCode::Unspecified;
    std::string operation, parameter, message, what;
    std::int64_t index = MV_ERROR_PARAMETER_INDEX_NONE;
  };
  std::shared_ptr<const Data> d_;
};

namespace detail {
// The error code a status implies when the C side left no matching record.
inline MvErrorCode implied_code(MvStatus s) noexcept {
  switch (s) {
    case MvStatus::InvalidArgument: return MvErrorCode::InvalidValue;
    case MvStatus::OutOfRange: return MvErrorCode::IndexOutOfRange;
    case MvStatus::TypeMismatch: return MvErrorCode::InvalidType;
    case MvStatus::NotFound: return MvErrorCode::MissingItem;
    case MvStatus::Unsupported: return MvErrorCode::UnsupportedValue;
    case MvStatus::InvalidState: return MvErrorCode::InvalidObjectState;
    case MvStatus::IoError: case MvStatus::DecodeError: case MvStatus::EncodeError: return MvErrorCode::IoFailure;
    case MvStatus::InternalError: return MvErrorCode::InternalFailure;
    default: return MvErrorCode::Unspecified;
  }
}

// Throws the C ABI's last error when `status` reports a failure.
inline void check(::MvStatus status, const char* op) {
  if (status >= 0) return;
  const MvStatus s = static_cast<MvStatus>(status);
  ::MvErrorInfo e{};
  if (mv_get_last_error_info(&e) == MV_STATUS_OK && e.status == status) {
    const ::MvStringView opname = e.operation.size ? e.operation : ::MvStringView{op, std::strlen(op)};
    throw MvException(s, static_cast<MvErrorCode>(e.code), opname, e.parameter, e.parameter_index, e.message);
  }
  throw MvException(s, implied_code(s), op, "", MV_ERROR_PARAMETER_INDEX_NONE, "operation failed");
}

[[noreturn]] inline void argument_error(MvErrorCode code, const char* op, const char* param, const char* msg) {
  throw MvException(MvStatus::InvalidArgument, code, op, param, MV_ERROR_PARAMETER_INDEX_NONE, msg);
}

// Owned C tuple released on scope exit.
struct OwnedTuple {
  ::MvTuple t = nullptr;
  OwnedTuple() = default;
  OwnedTuple(const OwnedTuple&) = delete;
  OwnedTuple& operator=(const OwnedTuple&) = delete;
  ~OwnedTuple() { if (t) mv_destroy_tuple(t); }
  ::MvTuple release() noexcept { return std::exchange(t, nullptr); }
};

// The empty control tuple passed for a default-constructed tuple or handle
// (one per thread; the C ABI only borrows inputs).
inline ::MvTuple empty_tuple() {
  thread_local OwnedTuple empty;
  if (!empty.t) check(mv_create_tuple(&empty.t, 0), "MvTuple");
  return empty.t;
}
}  // namespace detail

// ------------------------------------------------------------------- string
// include/mvkit/cpp/mv_string.hpp: UTF-8 text.
class MvString {
 public:
  MvString() noexcept = default;
  MvString(const char* value) : s_(value ? value : "") {}
  MvString(::MvStringView value) : s_(value.data ? value.data : "", value.data ? value.size : 0) {}
  MvString(std::string value) : s_(std::move(value)) {}
  const char* Text() const noexcept { return s_.c_str(); }
  std::size_t Length() const noexcept { return s_.size(); }
  ::MvStringView View() const noexcept { return {s_.data(), s_.size()}; }
  const std::string& Str() const noexcept { return s_; }
  bool operator==(const MvString& o) const noexcept { return s_ == o.s_; }
  bool operator!=(const MvString& o) const noexcept { return s_ != o.s_; }

 private:
  std::string s_;
};

class MvHandle;

// -------------------------------------------------------------------- tuple
// include/mvkit/cpp/mv_tuple.hpp: a control tuple of integers, reals, strings
// and handles.  A default tuple is empty and owns no C tuple until it is
// written or passed to the C ABI.  Copies are deep (handles in them share
// their models).
class MvTuple {
 public:
  MvTuple() noexcept = default;
  MvTuple(int value) : MvTuple(static_cast<std::int64_t>(value)) {}
  MvTuple(std::int64_t value) { Append(value); }
  MvTuple(double value) { Append(value); }
  MvTuple(const char* value) { Append(value); }
  MvTuple(const MvString& value) { Append(value); }
  inline MvTuple(const MvHandle& value);
  MvTuple(const std::vector<std::int64_t>& values) { for (auto v : values) Append(v); }
  MvTuple(const std::vector<double>& values) { for (auto v : values) Append(v); }
  MvTuple(const MvTuple& o) {
    if (o.t_) detail::check(mv_copy_tuple(o.t_, &t_), "MvTuple");
  }
  MvTuple(MvTuple&& o) noexcept : t_(std::exchange(o.t_, nullptr)) {}
  MvTuple& operator=(const MvTuple& o) {
    if (this != &o) {
      MvTuple copy(o);
      Swap(copy);
    }
    return *this;
  }
  MvTuple& operator=(MvTuple&& o) noexcept {
    if (this != &o) {
      MvTuple old(std::move(*this));
      t_ = std::exchange(o.t_, nullptr);
    }
    return *this;
  }
  ~MvTuple() noexcept { if (t_) mv_destroy_tuple(t_); }

  // Takes ownership of a C tuple (null: the empty tuple).
  static MvTuple Adopt(::MvTuple t) noexcept {
    MvTuple r;
    r.t_ = t;
    return r;
  }
  // The C tuple to pass as a borrowed control input.
  ::MvTuple ControlInput() const { return t_ ? t_ : detail::empty_tuple(); }

  std::size_t Length() const {
    if (!t_) return 0;
    std::size_t n = 0;
    detail::check(mv_length_tuple(t_, &n), "MvTuple");
    return n;
  }
  ::MvValueKind Kind(std::size_t index) const {
    ::MvValueKind k = MV_VALUE_EMPTY;
    detail::check(mv_get_type(ControlInput(), index, &k), "MvTuple");
    return k;
  }
  std::int64_t I(std::size_t index = 0) const {
    std::int64_t v = 0;
    detail::check(mv_get_i(ControlInput(), index, &v), "MvTuple");
    return v;
  }
  double D(std::size_t index = 0) const {
    double v = 0;
    detail::check(mv_get_d(ControlInput(), index, &v), "MvTuple");
    return v;
  }
  MvString S(std::size_t index = 0) const {
    ::MvStringView v{};
    detail::check(mv_get_s(ControlInput(), index, &v), "MvTuple");
    return MvString(v);
  }
  inline MvHandle H(std::size_t index = 0) const;

  MvTuple& Append(int value) { return Append(static_cast<std::int64_t>(value)); }
  MvTuple& Append(std::int64_t value) {
    const std::size_t n = Grow();
    detail::check(mv_set_i(t_, value, n), "MvTuple");
    return *this;
  }
  MvTuple& Append(double value) {
    const std::size_t n = Grow();
    detail::check(mv_set_d(t_, value, n), "MvTuple");
    return *this;
  }
  MvTuple& Append(const char* value) { return Append(MvString(value)); }
  MvTuple& Append(const MvString& value) {
    const std::size_t n = Grow();
    detail::check(mv_set_s(t_, value.View(), n), "MvTuple");
    return *this;
  }
  inline MvTuple& Append(const MvHandle& value);

  // Elements of this tuple followed by those of `other`.
  MvTuple TupleConcat(const MvTuple& other) const {
    MvTuple r(*this);
    for (std::size_t i = 0, n = other.Length(); i < n; ++i) r.AppendFrom(other, i);
    return r;
  }
  // The elements at the integer positions in `indices`, in that order.
  MvTuple TupleSelect(const MvTuple& indices) const {
    MvTuple r;
    const std::size_t n = Length();
    for (std::size_t k = 0, m = indices.Length(); k < m; ++k) {
      const std::int64_t i = indices.I(k);
      if (i < 0 || static_cast<std::size_t>(i) >= n)
        throw MvException(MvStatus::OutOfRange, MvErrorCode::IndexOutOfRange, "TupleSelect", "Index",
                          static_cast<std::int64_t>(k), "index outside the tuple");
      r.AppendFrom(*this, static_cast<std::size_t>(i));
    }
    return r;
  }
  // The elements from position `left` through `right` (inclusive).
  MvTuple TupleSelectRange(const MvTuple& left, const MvTuple& right) const {
    const std::int64_t a = left.I(), b = right.I(), n = static_cast<std::int64_t>(Length());
    if (a < 0 || b >= n || (a > b + 1))
      throw MvException(MvStatus::OutOfRange, MvErrorCode::IndexOutOfRange, "TupleSelectRange", "Leftindex",
                        MV_ERROR_PARAMETER_INDEX_NONE, "range outside the tuple");
    MvTuple r;
    for (std::int64_t i = a; i <= b; ++i) r.AppendFrom(*this, static_cast<std::size_t>(i));
    return r;
  }

  void Swap(MvTuple& o) noexcept { std::swap(t_, o.t_); }

 private:
  std::size_t Grow() {
    if (!t_) detail::check(mv_create_tuple(&t_, 0), "MvTuple");
    const std::size_t n = Length();
    detail::check(mv_resize_tuple(t_, n + 1), "MvTuple");
    return n;
  }
  void AppendFrom(const MvTuple& src, std::size_t i);

  ::MvTuple t_ = nullptr;
};

// ------------------------------------------------------------------- handle
// include/mvkit/cpp/mv_handle.hpp: a shared reference to a control object,
// kept as a one-element C tuple holding the handle.
class MvSerializedItem;
class MvHandle {
 public:
  MvHandle() noexcept = default;
  // A new reference to the object `handle` refers to.
  static MvHandle FromNative(::MvHandle handle) {
    if (!handle) detail::argument_error(MvErrorCode::NullArgument, "MvHandle", "handle", "handle must not be null");
    MvHandle h;
    detail::check(mv_create_tuple(&h.holder_.t, 1), "MvHandle");
    detail::check(mv_set_h(h.holder_.t, handle, 0), "MvHandle");
    return h;
  }
  MvHandle(const MvHandle& o) { if (o.holder_.t) detail::check(mv_copy_tuple(o.holder_.t, &holder_.t), "MvHandle"); }
  MvHandle(MvHandle&& o) noexcept { holder_.t = o.holder_.release(); }
  MvHandle& operator=(const MvHandle& o) {
    if (this != &o) {
      MvHandle copy(o);
      std::swap(holder_.t, copy.holder_.t);
    }
    return *this;
  }
  MvHandle& operator=(MvHandle&& o) noexcept {
    if (this != &o) {
      MvHandle old(std::move(*this));
      holder_.t = o.holder_.release();
    }
    return *this;
  }
  virtual ~MvHandle() noexcept = default;

  // Whether this refers to an object that has not been cleared.
  bool IsInitialized() const noexcept {
    try {
      return Type() != MvObjectType::Unknown;
    } catch (...) {
      return false;
    }
  }
  MvObjectType Type() const {
    if (!holder_.t) return MvObjectType::Unknown;
    detail::OwnedTuple sem;
    detail::check(mv_tuple_sem_type(holder_.t, &sem.t), "MvHandle");
    ::MvStringView v{};
    detail::check(mv_get_s(sem.t, 0, &v), "MvHandle");
    const std::string name(v.data, v.size);
    if (name == "ncc_model") return MvObjectType::NccModel;
    if (name == "serialized_item") return MvObjectType::SerializedItem;
    return MvObjectType::Unknown;
  }
  ::MvHandle NativeHandle() const noexcept {
    ::MvHandle h = nullptr;
    if (holder_.t && mv_get_h(holder_.t, 0, &h) != MV_STATUS_OK) h = nullptr;
    return h;
  }
  // The control tuple holding this handle (borrowed; owned by this object).
  ::MvTuple ControlInput() const { return holder_.t ? holder_.t : detail::empty_tuple(); }
  inline MvSerializedItem SerializeHandle() const;
  inline void DeserializeHandle(const MvSerializedItem& item);

 protected:
  // Replaces this reference with the first handle of `t`.
  void Assign(const MvTuple& t) {
    MvHandle h = t.H(0);
    *this = std::move(h);
  }

 private:
  friend class MvTuple;
  detail::OwnedTuple holder_;
};

inline MvTuple::MvTuple(const MvHandle& value) { Append(value); }
inline MvHandle MvTuple::H(std::size_t index) const {
  ::MvHandle h = nullptr;
  detail::check(mv_get_h(ControlInput(), index, &h), "MvTuple");
  return MvHandle::FromNative(h);
}
inline MvTuple& MvTuple::Append(const MvHandle& value) {
  const ::MvHandle h = value.NativeHandle();
  if (!h) detail::argument_error(MvErrorCode::NullArgument, "MvTuple", "value", "uninitialized handle");
  const std::size_t n = Grow();
  detail::check(mv_set_h(t_, h, n), "MvTuple");
  return *this;
}
inline void MvTuple::AppendFrom(const MvTuple& src, std::size_t i) {
  switch (src.Kind(i)) {
    case MV_VALUE_INT64: Append(src.I(i)); break;
    case MV_VALUE_REAL64: Append(src.D(i)); break;
    case MV_VALUE_STRING: Append(src.S(i)); break;
    case MV_VALUE_HANDLE: Append(src.H(i)); break;
    default: Grow(); break;  // an empty slot stays empty
  }
}

// Clears every handle in `Handle`; all references to those objects see the
// cleared state (include/mvkit/c/mv_handle.h:40).
inline void ClearHandle(const MvTuple& Handle) { detail::check(mv_clear_handle(Handle.ControlInput()), "ClearHandle"); }
inline void TupleSemType(const MvTuple& T, MvTuple* SemType) {
  if (!SemType) detail::argument_error(MvErrorCode::NullArgument, "TupleSemType", "output", "output must not be null");
  detail::OwnedTuple out;
  detail::check(mv_tuple_sem_type(T.ControlInput(), &out.t), "TupleSemType");
  *SemType = MvTuple::Adopt(out.release());
}

namespace detail {
// Owned outputs of one C call, committed to the caller's tuples together.
template <std::size_t N>
class Outputs {
 public:
  Outputs(const char* op, std::initializer_list<MvTuple*> dest) {
    std::size_t i = 0;
    for (MvTuple* d : dest) {
      if (!d) argument_error(MvErrorCode::NullArgument, op, "output", "output must not be null");
      for (std::size_t j = 0; j < i; ++j)
        if (dest_[j] == d) argument_error(MvErrorCode::InvalidValue, op, "Output", "output pointers must be distinct");
      dest_[i++] = d;
    }
  }
  Outputs(const Outputs&) = delete;
  Outputs& operator=(const Outputs&) = delete;
  ~Outputs() { for (::MvTuple t : raw_) if (t) mv_destroy_tuple(t); }
  ::MvTuple* operator[](std::size_t i) noexcept { return &raw_[i]; }
  void commit() noexcept {
    for (std::size_t i = 0; i < N; ++i) *dest_[i] = MvTuple::Adopt(std::exchange(raw_[i], nullptr));
  }

 private:
  MvTuple* dest_[N] = {};
  ::MvTuple raw_[N] = {};
};
}  // namespace detail

inline void SerializeHandle(const MvTuple& Handle, MvTuple* SerializedItem) {
  detail::Outputs<1> out("SerializeHandle", {SerializedItem});
  detail::check(mv_serialize_handle(Handle.ControlInput(), out[0]), "SerializeHandle");
  out.commit();
}
inline void DeserializeHandle(const MvTuple& SerializedItem, MvTuple* Handle) {
  detail::Outputs<1> out("DeserializeHandle", {Handle});
  detail::check(mv_deserialize_handle(SerializedItem.ControlInput(), out[0]), "DeserializeHandle");
  out.commit();
}
inline void CreateSerializedItemPtr(const MvTuple& Pointer, const MvTuple& Size, const MvTuple& Copy,
                                    MvTuple* SerializedItemHandle) {
  detail::Outputs<1> out("CreateSerializedItemPtr", {SerializedItemHandle});
  detail::check(mv_create_serialized_item_ptr(Pointer.ControlInput(), Size.ControlInput(), Copy.ControlInput(), out[0]),
                "CreateSerializedItemPtr");
  out.commit();
}
inline void GetSerializedItemPtr(const MvTuple& SerializedItemHandle, MvTuple* Pointer, MvTuple* Size) {
  detail::Outputs<2> out("GetSerializedItemPtr", {Pointer, Size});
  detail::check(mv_get_serialized_item_ptr(SerializedItemHandle.ControlInput(), out[0], out[1]),
                "GetSerializedItemPtr");
  out.commit();
}

// include/mvkit/cpp/mv_handle.hpp MvSerializedItem: the bytes of a
// serialized control object.
class MvSerializedItem : public MvHandle {
 public:
  MvSerializedItem() noexcept = default;
  explicit MvSerializedItem(const MvHandle& handle) : MvHandle(handle) {
    if (Type() != MvObjectType::SerializedItem)
      throw MvException(MvStatus::TypeMismatch, MvErrorCode::InvalidType, "MvSerializedItem", "SerializedItem",
                        MV_ERROR_PARAMETER_INDEX_NONE, "control handle is not a serialized item");
  }
  MvSerializedItem(void* Pointer, std::int64_t Size, const char* Copy) { CreateSerializedItemPtr(Pointer, Size, Copy); }
  MvSerializedItem(void* Pointer, std::int64_t Size, const MvString& Copy) {
    CreateSerializedItemPtr(Pointer, Size, Copy);
  }
  void CreateSerializedItemPtr(void* Pointer, std::int64_t Size, const MvString& Copy) {
    MvTuple item;
    mv::CreateSerializedItemPtr(MvTuple(static_cast<std::int64_t>(reinterpret_cast<std::uintptr_t>(Pointer))),
                                MvTuple(Size), MvTuple(Copy), &item);
    Assign(item);
  }
  void CreateSerializedItemPtr(void* Pointer, std::int64_t Size, const char* Copy) {
    CreateSerializedItemPtr(Pointer, Size, MvString(Copy));
  }
  // The item's bytes (valid while the item is) and their count.
  void* GetSerializedItemPtr(std::int64_t* Size) const {
    if (!Size) detail::argument_error(MvErrorCode::NullArgument, "GetSerializedItemPtr", "output", "output must not be null");
    MvTuple p, n;
    mv::GetSerializedItemPtr(MvTuple(*this), &p, &n);
    const std::int64_t address = p.I(), count = n.I();
    *Size = count;
    return reinterpret_cast<void*>(static_cast<std::uintptr_t>(address));
  }
};

inline MvSerializedItem MvHandle::SerializeHandle() const {
  MvTuple item;
  mv::SerializeHandle(MvTuple(*this), &item);
  return MvSerializedItem(item.H(0));
}
inline void MvHandle::DeserializeHandle(const MvSerializedItem& item) {
  MvTuple h;
  mv::DeserializeHandle(MvTuple(item), &h);
  Assign(h);
}

// ---------------------------------------------------------- iconic objects
// include/mvkit/cpp/mv_iconic.hpp: an image or region array, shared by copies.
class MvObject {
 public:
  MvObject() { GenEmptyObj(); }
  virtual ~MvObject() noexcept = default;
  MvObject(const MvObject&) = default;
  MvObject(MvObject&&) noexcept = default;
  MvObject& operator=(const MvObject&) = default;
  MvObject& operator=(MvObject&&) noexcept = default;

  // Takes ownership of a C object.
  static MvObject FromNative(::MvObject handle) {
    if (!handle) detail::argument_error(MvErrorCode::NullArgument, "MvObject", "handle", "object must not be null");
    MvObject o(handle, Adopt{});
    return o;
  }
  void GenEmptyObj() {
    ::MvObject h = nullptr;
    detail::check(mv_gen_empty_obj(&h), "GenEmptyObj");
    Reset(h);
  }
  ::MvObject NativeHandle() const noexcept { return obj_.get(); }

 protected:
  struct Adopt {};
  MvObject(::MvObject handle, Adopt) { Reset(handle); }
  void Reset(::MvObject handle) { obj_.reset(handle, [](::MvObject h) { mv_clear_obj(h); }); }

 private:
  std::shared_ptr<std::remove_pointer<::MvObject>::type> obj_;
};

inline void GenEmptyObj(MvObject* objects) {
  if (!objects) detail::argument_error(MvErrorCode::NullArgument, "GenEmptyObj", "output", "output must not be null");
  objects->GenEmptyObj();
}

// include/mvkit/cpp/mv_region.hpp (the parts a reduced-domain search uses)
class MvRegion : public MvObject {
 public:
  MvRegion() = default;
  explicit MvRegion(const MvObject& o) : MvObject(o) {}
  void GenRectangle1(const MvTuple& Row1, const MvTuple& Column1, const MvTuple& Row2, const MvTuple& Column2) {
    ::MvObject h = nullptr;
    detail::check(mv_gen_rectangle1(&h, Row1.ControlInput(), Column1.ControlInput(), Row2.ControlInput(),
                                    Column2.ControlInput()),
                  "GenRectangle1");
    Reset(h);
  }
  void GenRegionRuns(const MvTuple& Row, const MvTuple& ColumnBegin, const MvTuple& ColumnEnd) {
    ::MvObject h = nullptr;
    detail::check(mv_gen_region_runs(&h, Row.ControlInput(), ColumnBegin.ControlInput(), ColumnEnd.ControlInput()),
                  "GenRegionRuns");
    Reset(h);
  }
  void GetRegionRuns(MvTuple* Row, MvTuple* ColumnBegin, MvTuple* ColumnEnd) const {
    detail::Outputs<3> out("GetRegionRuns", {Row, ColumnBegin, ColumnEnd});
    detail::check(mv_get_region_runs(NativeHandle(), out[0], out[1], out[2]), "GetRegionRuns");
    out.commit();
  }

 private:
  friend class MvImage;
  MvRegion(::MvObject handle, Adopt a) : MvObject(handle, a) {}
};

class MvNCCModel;
class MvNCCModelArray;

// include/mvkit/cpp/mv_image.hpp (construction, domains and the NCC members)
class MvImage : public MvObject {
 public:
  MvImage() = default;
  explicit MvImage(const MvObject& o) : MvObject(o) {}
  // Copies a row-major byte or uint2 plane (stride == width).
  void GenImage1(const MvTuple& Type, const MvTuple& Width, const MvTuple& Height, const MvTuple& PixelPointer) {
    ::MvObject h = nullptr;
    detail::check(mv_gen_image1(&h, Type.ControlInput(), Width.ControlInput(), Height.ControlInput(),
                                PixelPointer.ControlInput()),
                  "GenImage1");
    Reset(h);
  }
  void GenImage1(const char* Type, std::int64_t Width, std::int64_t Height, const void* PixelPointer) {
    GenImage1(MvTuple(Type), MvTuple(Width), MvTuple(Height), MvTuple(address(PixelPointer)));
  }
  // Borrows the plane; `ClearProc(pointer)` is called when the last image
  // sharing it is released.
  void GenImage1Extern(const char* Type, std::int64_t Width, std::int64_t Height, const void* PixelPointer,
                       void (*ClearProc)(void*)) {
    ::MvObject h = nullptr;
    const MvTuple t(Type), w(Width), hh(Height), p(address(PixelPointer)),
        c(static_cast<std::int64_t>(reinterpret_cast<std::uintptr_t>(ClearProc)));
    detail::check(mv_gen_image1_extern(&h, t.ControlInput(), w.ControlInput(), hh.ControlInput(), p.ControlInput(),
                                       c.ControlInput()),
                  "GenImage1Extern");
    Reset(h);
  }
  MvImage ReduceDomain(const MvRegion& Region) const {
    ::MvObject h = nullptr;
    detail::check(mv_reduce_domain(NativeHandle(), Region.NativeHandle(), &h), "ReduceDomain");
    return MvImage(h, Adopt{});
  }
  MvImage FullDomain() const {
    ::MvObject h = nullptr;
    detail::check(mv_full_domain(NativeHandle(), &h), "FullDomain");
    return MvImage(h, Adopt{});
  }
  MvRegion GetDomain() const {
    ::MvObject h = nullptr;
    detail::check(mv_get_domain(NativeHandle(), &h), "GetDomain");
    return MvRegion(h, Adopt{});
  }

  // mv_image.hpp's NCC members: this image as template or search image.
  inline MvNCCModel CreateNccModel(const MvTuple& NumLevels, double AngleStart, double AngleExtent,
                                   const MvTuple& AngleStep, const MvString& Metric) const;
  inline MvNCCModel CreateNccModel(const MvTuple& NumLevels, double AngleStart, double AngleExtent,
                                   const MvTuple& AngleStep, const char* Metric) const;
  inline MvNCCModel CreateNccModel(std::int64_t NumLevels, double AngleStart, double AngleExtent, double AngleStep,
                                   const MvString& Metric) const;
  inline MvNCCModel CreateNccModel(std::int64_t NumLevels, double AngleStart, double AngleExtent, double AngleStep,
                                   const char* Metric) const;
  inline void FindNccModel(const MvNCCModel& ModelID, double AngleStart, double AngleExtent, double MinScore,
                           std::int64_t NumMatches, double MaxOverlap, const MvString& SubPixel,
                           const MvTuple& NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle,
                           MvTuple* Score) const;
  inline void FindNccModel(const MvNCCModel& ModelID, double AngleStart, double AngleExtent, double MinScore,
                           std::int64_t NumMatches, double MaxOverlap, const MvString& SubPixel,
                           std::int64_t NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle,
                           MvTuple* Score) const;
  inline void FindNccModel(const MvNCCModel& ModelID, double AngleStart, double AngleExtent, double MinScore,
                           std::int64_t NumMatches, double MaxOverlap, const char* SubPixel, std::int64_t NumLevels,
                           MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score) const;
  inline void FindNccModels(const MvNCCModelArray& ModelIDs, const MvTuple& AngleStart, const MvTuple& AngleExtent,
                            const MvTuple& MinScore, const MvTuple& NumMatches, const MvTuple& MaxOverlap,
                            const MvTuple& SubPixel, const MvTuple& NumLevels, MvTuple* Row, MvTuple* Column,
                            MvTuple* Angle, MvTuple* Score, MvTuple* Model) const;

 private:
  static std::int64_t address(const void* p) noexcept {
    return static_cast<std::int64_t>(reinterpret_cast<std::uintptr_t>(p));
  }
  MvImage(::MvObject handle, Adopt a) : MvObject(handle, a) {}
};

inline void GenImage1(MvObject* Image, const MvTuple& Type, const MvTuple& Width, const MvTuple& Height,
                      const MvTuple& PixelPointer) {
  if (!Image) detail::argument_error(MvErrorCode::NullArgument, "GenImage1", "output", "output must not be null");
  MvImage im;
  im.GenImage1(Type, Width, Height, PixelPointer);
  *Image = im;
}
inline void ReduceDomain(const MvObject& Image, const MvObject& Region, MvObject* ImageReduced) {
  if (!ImageReduced)
    detail::argument_error(MvErrorCode::NullArgument, "ReduceDomain", "output", "output must not be null");
  *ImageReduced = MvImage(Image).ReduceDomain(MvRegion(Region));
}

// ------------------------------------------------- NCC operators (procedural)
// include/mvkit/cpp/mv_ncc_matching.hpp:450-482
inline void CreateNccModel(const MvObject& Template, const MvTuple& NumLevels, const MvTuple& AngleStart,
                           const MvTuple& AngleExtent, const MvTuple& AngleStep, const MvTuple& Metric,
                           MvTuple* ModelID) {
  detail::Outputs<1> out("CreateNccModel", {ModelID});
  detail::check(mv_create_ncc_model(Template.NativeHandle(), NumLevels.ControlInput(), AngleStart.ControlInput(),
                                    AngleExtent.ControlInput(), AngleStep.ControlInput(), Metric.ControlInput(),
                                    out[0]),
                "CreateNccModel");
  out.commit();
}
inline void GetNccModelParams(const MvTuple& ModelID, MvTuple* NumLevels, MvTuple* AngleStart, MvTuple* AngleExtent,
                              MvTuple* AngleStep, MvTuple* Metric) {
  detail::Outputs<5> out("GetNccModelParams", {NumLevels, AngleStart, AngleExtent, AngleStep, Metric});
  detail::check(mv_get_ncc_model_params(ModelID.ControlInput(), out[0], out[1], out[2], out[3], out[4]),
                "GetNccModelParams");
  out.commit();
}
inline void FindNccModel(const MvObject& Image, const MvTuple& ModelID, const MvTuple& AngleStart,
                         const MvTuple& AngleExtent, const MvTuple& MinScore, const MvTuple& NumMatches,
                         const MvTuple& MaxOverlap, const MvTuple& SubPixel, const MvTuple& NumLevels, MvTuple* Row,
                         MvTuple* Column, MvTuple* Angle, MvTuple* Score) {
  detail::Outputs<4> out("FindNccModel", {Row, Column, Angle, Score});
  detail::check(mv_find_ncc_model(Image.NativeHandle(), ModelID.ControlInput(), AngleStart.ControlInput(),
                                  AngleExtent.ControlInput(), MinScore.ControlInput(), NumMatches.ControlInput(),
                                  MaxOverlap.ControlInput(), SubPixel.ControlInput(), NumLevels.ControlInput(),
                                  out[0], out[1], out[2], out[3]),
                "FindNccModel");
  out.commit();
}
inline void FindNccModels(const MvObject& Image, const MvTuple& ModelIDs, const MvTuple& AngleStart,
                          const MvTuple& AngleExtent, const MvTuple& MinScore, const MvTuple& NumMatches,
                          const MvTuple& MaxOverlap, const MvTuple& SubPixel, const MvTuple& NumLevels, MvTuple* Row,
                          MvTuple* Column, MvTuple* Angle, MvTuple* Score, MvTuple* Model) {
  detail::Outputs<5> out("FindNccModels", {Row, Column, Angle, Score, Model});
  detail::check(mv_find_ncc_models(Image.NativeHandle(), ModelIDs.ControlInput(), AngleStart.ControlInput(),
                                   AngleExtent.ControlInput(), MinScore.ControlInput(), NumMatches.ControlInput(),
                                   MaxOverlap.ControlInput(), SubPixel.ControlInput(), NumLevels.ControlInput(),
                                   out[0], out[1], out[2], out[3], out[4]),
                "FindNccModels");
  out.commit();
}
inline void ReadNccModel(const MvTuple& FileName, MvTuple* ModelID) {
  detail::Outputs<1> out("ReadNccModel", {ModelID});
  detail::check(mv_read_ncc_model(FileName.ControlInput(), out[0]), "ReadNccModel");
  out.commit();
}
inline void WriteNccModel(const MvTuple& ModelID, const MvTuple& FileName) {
  detail::check(mv_write_ncc_model(ModelID.ControlInput(), FileName.ControlInput()), "WriteNccModel");
}

// ---------------------------------------------------- NCC model (object API)
// include/mvkit/cpp/mv_ncc_matching.hpp:10-257.  Copies share one model.
class MvNCCModel : public MvHandle {
 public:
  MvNCCModel() noexcept = default;
  // A typed reference to `handle`, which must be a live NCC model.
  explicit MvNCCModel(const MvHandle& handle) : MvHandle(handle) {
    if (Type() != MvObjectType::NccModel)
      throw MvException(MvStatus::TypeMismatch, MvErrorCode::InvalidType, "MvNCCModel", "ModelID",
                        MV_ERROR_PARAMETER_INDEX_NONE, "control handle is not an NCC model");
  }

  void CreateNccModel(const MvImage& Template, const MvTuple& NumLevels, double AngleStart, double AngleExtent,
                      const MvTuple& AngleStep, const MvString& Metric) {
    MvTuple id;
    mv::CreateNccModel(Template, NumLevels, AngleStart, AngleExtent, AngleStep, Metric, &id);
    Assign(id);
  }
  void CreateNccModel(const MvImage& Template, const MvTuple& NumLevels, double AngleStart, double AngleExtent,
                      const MvTuple& AngleStep, const char* Metric) {
    CreateNccModel(Template, NumLevels, AngleStart, AngleExtent, AngleStep, MvString(Metric));
  }
  void CreateNccModel(const MvImage& Template, std::int64_t NumLevels, double AngleStart, double AngleExtent,
                      double AngleStep, const MvString& Metric) {
    CreateNccModel(Template, MvTuple(NumLevels), AngleStart, AngleExtent, MvTuple(AngleStep), Metric);
  }
  void CreateNccModel(const MvImage& Template, std::int64_t NumLevels, double AngleStart, double AngleExtent,
                      double AngleStep, const char* Metric) {
    CreateNccModel(Template, MvTuple(NumLevels), AngleStart, AngleExtent, MvTuple(AngleStep), MvString(Metric));
  }

  // The model's training parameters; returns the number of levels.  The
  // outputs are written only when every one of them was read.
  std::int64_t GetNccModelParams(double* AngleStart, double* AngleExtent, double* AngleStep, MvString* Metric) const {
    constexpr const char* op = "GetNccModelParams";
    if (!AngleStart || !AngleExtent || !AngleStep || !Metric)
      detail::argument_error(MvErrorCode::NullArgument, op, "output", "output must not be null");
    const void* p[4] = {AngleStart, AngleExtent, AngleStep, Metric};
    const std::size_t n[4] = {sizeof(double), sizeof(double), sizeof(double), sizeof(MvString)};
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < i; ++j) {
        const auto a = reinterpret_cast<std::uintptr_t>(p[i]), b = reinterpret_cast<std::uintptr_t>(p[j]);
        if (a < b + n[j] && b < a + n[i]) detail::argument_error(MvErrorCode::InvalidValue, op, "Output",
                                                                 "output ranges must be distinct");
      }
    MvTuple levels, start, extent, step, metric;
    mv::GetNccModelParams(MvTuple(*this), &levels, &start, &extent, &step, &metric);
    const std::int64_t l = levels.I();
    const double s = start.D(), e = extent.D(), st = step.D();
    MvString m = metric.S();
    *AngleStart = s;
    *AngleExtent = e;
    *AngleStep = st;
    *Metric = std::move(m);
    return l;
  }

  void FindNccModel(const MvImage& Image, double AngleStart, double AngleExtent, double MinScore,
                    std::int64_t NumMatches, double MaxOverlap, const MvString& SubPixel, const MvTuple& NumLevels,
                    MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score) const {
    mv::FindNccModel(Image, MvTuple(*this), AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, SubPixel,
                     NumLevels, Row, Column, Angle, Score);
  }
  void FindNccModel(const MvImage& Image, double AngleStart, double AngleExtent, double MinScore,
                    std::int64_t NumMatches, double MaxOverlap, const MvString& SubPixel, std::int64_t NumLevels,
                    MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score) const {
    FindNccModel(Image, AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, SubPixel, MvTuple(NumLevels), Row,
                 Column, Angle, Score);
  }
  void FindNccModel(const MvImage& Image, double AngleStart, double AngleExtent, double MinScore,
                    std::int64_t NumMatches, double MaxOverlap, const char* SubPixel, std::int64_t NumLevels,
                    MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score) const {
    FindNccModel(Image, AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, MvString(SubPixel),
                 MvTuple(NumLevels), Row, Column, Angle, Score);
  }

  inline static void FindNccModels(const MvImage& Image, const MvNCCModelArray& ModelIDs, const MvTuple& AngleStart,
                                   const MvTuple& AngleExtent, const MvTuple& MinScore, const MvTuple& NumMatches,
                                   const MvTuple& MaxOverlap, const MvTuple& SubPixel, const MvTuple& NumLevels,
                                   MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score, MvTuple* Model);
  // This model alone through find_models (Model receives 0 per match).
  void FindNccModels(const MvImage& Image, double AngleStart, double AngleExtent, double MinScore,
                     std::int64_t NumMatches, double MaxOverlap, const MvString& SubPixel, std::int64_t NumLevels,
                     MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score, MvTuple* Model) const {
    mv::FindNccModels(Image, MvTuple(*this), AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, SubPixel,
                      NumLevels, Row, Column, Angle, Score, Model);
  }
  void FindNccModels(const MvImage& Image, double AngleStart, double AngleExtent, double MinScore,
                     std::int64_t NumMatches, double MaxOverlap, const char* SubPixel, std::int64_t NumLevels,
                     MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score, MvTuple* Model) const {
    FindNccModels(Image, AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, MvString(SubPixel), NumLevels,
                  Row, Column, Angle, Score, Model);
  }

  void ReadNccModel(const MvString& FileName) {
    MvTuple id;
    mv::ReadNccModel(FileName, &id);
    Assign(id);
  }
  void ReadNccModel(const char* FileName) { ReadNccModel(MvString(FileName)); }
  void WriteNccModel(const MvString& FileName) const { mv::WriteNccModel(MvTuple(*this), FileName); }
  void WriteNccModel(const char* FileName) const { WriteNccModel(MvString(FileName)); }
};

// include/mvkit/cpp/mv_ncc_matching.hpp:262-380: an array of models shared
// by copies.
class MvNCCModelArray {
 public:
  MvNCCModelArray() noexcept = default;
  MvNCCModelArray(const MvNCCModel* models, std::int64_t length) {
    if (length < 0 || (length > 0 && !models))
      throw MvException(MvStatus::InvalidArgument, MvErrorCode::InvalidValue, "MvNCCModelArray", "models",
                        MV_ERROR_PARAMETER_INDEX_NONE, "nonnegative length and valid model array required");
    if (length) models_ = std::make_shared<const std::vector<MvNCCModel>>(models, models + length);
  }
  virtual ~MvNCCModelArray() noexcept = default;
  MvNCCModelArray(const MvNCCModelArray&) = default;
  MvNCCModelArray(MvNCCModelArray&&) noexcept = default;
  MvNCCModelArray& operator=(const MvNCCModelArray&) = default;
  MvNCCModelArray& operator=(MvNCCModelArray&&) noexcept = default;

  virtual void Clear() noexcept { models_.reset(); }
  const MvNCCModel* Tools() const noexcept { return models_ ? models_->data() : nullptr; }
  virtual std::int64_t Length() const noexcept { return models_ ? std::int64_t(models_->size()) : 0; }
  // Every element of `handles` must be a live NCC model.
  virtual void SetFromTuple(const MvTuple& handles) {
    const std::size_t n = handles.Length();
    if (!n) {
      Clear();
      return;
    }
    auto next = std::make_shared<std::vector<MvNCCModel>>();
    next->reserve(n);
    for (std::size_t i = 0; i < n; ++i) next->emplace_back(handles.H(i));
    models_ = std::move(next);
  }
  virtual MvTuple ConvertToTuple() const {
    MvTuple t;
    if (models_)
      for (const MvNCCModel& m : *models_) t.Append(m);
    return t;
  }

 private:
  std::shared_ptr<const std::vector<MvNCCModel>> models_;
};

inline void MvNCCModel::FindNccModels(const MvImage& Image, const MvNCCModelArray& ModelIDs,
                                      const MvTuple& AngleStart, const MvTuple& AngleExtent, const MvTuple& MinScore,
                                      const MvTuple& NumMatches, const MvTuple& MaxOverlap, const MvTuple& SubPixel,
                                      const MvTuple& NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle,
                                      MvTuple* Score, MvTuple* Model) {
  mv::FindNccModels(Image, ModelIDs.ConvertToTuple(), AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap,
                    SubPixel, NumLevels, Row, Column, Angle, Score, Model);
}

inline MvNCCModel MvImage::CreateNccModel(const MvTuple& NumLevels, double AngleStart, double AngleExtent,
                                          const MvTuple& AngleStep, const MvString& Metric) const {
  MvNCCModel m;
  m.CreateNccModel(*this, NumLevels, AngleStart, AngleExtent, AngleStep, Metric);
  return m;
}
inline MvNCCModel MvImage::CreateNccModel(const MvTuple& NumLevels, double AngleStart, double AngleExtent,
                                          const MvTuple& AngleStep, const char* Metric) const {
  return CreateNccModel(NumLevels, AngleStart, AngleExtent, AngleStep, MvString(Metric));
}
inline MvNCCModel MvImage::CreateNccModel(std::int64_t NumLevels, double AngleStart, double AngleExtent,
                                          double AngleStep, const MvString& Metric) const {
  return CreateNccModel(MvTuple(NumLevels), AngleStart, AngleExtent, MvTuple(AngleStep), Metric);
}
inline MvNCCModel MvImage::CreateNccModel(std::int64_t NumLevels, double AngleStart, double AngleExtent,
                                          double AngleStep, const char* Metric) const {
  return CreateNccModel(MvTuple(NumLevels), AngleStart, AngleExtent, MvTuple(AngleStep), MvString(Metric));
}
inline void MvImage::FindNccModel(const MvNCCModel& ModelID, double AngleStart, double AngleExtent, double MinScore,
                                  std::int64_t NumMatches, double MaxOverlap, const MvString& SubPixel,
                                  const MvTuple& NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle,
                                  MvTuple* Score) const {
  ModelID.FindNccModel(*this, AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, SubPixel, NumLevels, Row,
                       Column, Angle, Score);
}
inline void MvImage::FindNccModel(const MvNCCModel& ModelID, double AngleStart, double AngleExtent, double MinScore,
                                  std::int64_t NumMatches, double MaxOverlap, const MvString& SubPixel,
                                  std::int64_t NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle,
                                  MvTuple* Score) const {
  ModelID.FindNccModel(*this, AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, SubPixel, NumLevels, Row,
                       Column, Angle, Score);
}
inline void MvImage::FindNccModel(const MvNCCModel& ModelID, double AngleStart, double AngleExtent, double MinScore,
                                  std::int64_t NumMatches, double MaxOverlap, const char* SubPixel,
                                  std::int64_t NumLevels, MvTuple* Row, MvTuple* Column, MvTuple* Angle,
                                  MvTuple* Score) const {
  ModelID.FindNccModel(*this, AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, SubPixel, NumLevels, Row,
                       Column, Angle, Score);
}
inline void MvImage::FindNccModels(const MvNCCModelArray& ModelIDs, const MvTuple& AngleStart,
                                   const MvTuple& AngleExtent, const MvTuple& MinScore, const MvTuple& NumMatches,
                                   const MvTuple& MaxOverlap, const MvTuple& SubPixel, const MvTuple& NumLevels,
                                   MvTuple* Row, MvTuple* Column, MvTuple* Angle, MvTuple* Score,
                                   MvTuple* Model) const {
  MvNCCModel::FindNccModels(*this, ModelIDs, AngleStart, AngleExtent, MinScore, NumMatches, MaxOverlap, SubPixel,
                            NumLevels, Row, Column, Angle, Score, Model);
}

}  // namespace mv

#endif
