// Source-compatible subset of the reference's C++ host interface (Model,
// Group, field spans, archives, Neighborhood, SchemeBase, SchemeRegistry),
// with GPU schemes that run every procedure through the C ABI of
// include/prtcl_hip.h.
//
// Mirrors: src/prtcl/data/model.hpp:24-120 (Model), data/group.hpp:28-150
// (Group), data/varying_field.hpp:25-58 and uniform_field.hpp:21-53 (spans),
// data/component_type.cpp, data/shape.cpp and data/tensor_type.cpp (field
// types), util/archive.hpp:13-73 (archives) with the Save/Load members of
// model.cpp:54-75, group.cpp:31-59, varying_manager.cpp:21-45 and
// uniform_manager.cpp:22-41, util/neighborhood.hpp:23-42 (Neighborhood),
// schemes/scheme_base.hpp:18-114 (SchemeBase, SchemeRegistry,
// SchemeRegistration). Differences: spans index
// to a pointer to the item's components (no Eigen maps); range accessors
// return std::vector copies of the names; errors are the reference's string
// literals where it throws them ("procedure does not exist", "scheme is not
// registered") and typed exceptions otherwise.
//
// GPU schemes are registered under the reference names with ", B=hip"
// appended, e.g. "prtcl::schemes::iisph[T=f32, N=3, K=Cubic Spline, B=hip]".
// A model's groups of type "fluid" and "boundary" are mirrored by one device
// scene; every procedure uploads the mirrored fields present on the host,
// runs, and downloads them again, so host code observes the field
// transitions of the CPU schemes.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <istream>
#include <limits>
#include <map>
#include <memory>
#include <ostream>
#include <set>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <vector>

#include "../../include/prtcl_hip.h"

namespace prtcl {

struct FieldOfDifferentTypeAlreadyExistsError : std::logic_error {
  FieldOfDifferentTypeAlreadyExistsError()
      : std::logic_error("field of different type already exists") {}
};
struct FieldOfDifferentKindAlreadyExistsError : std::logic_error {
  FieldOfDifferentKindAlreadyExistsError()
      : std::logic_error("field of different kind already exists") {}
};
struct GroupOfDifferentTypeAlreadyExists : std::logic_error {
  GroupOfDifferentTypeAlreadyExists()
      : std::logic_error("group of different type already exists") {}
};

enum class GroupIndex : size_t { kNoIndex = std::numeric_limits<size_t>::max() };

// ---- archives (util/archive.hpp:13-73) ------------------------------------
class ArchiveWriter {
 public:
  virtual ~ArchiveWriter() = default;
  virtual void SaveString(std::string_view value) = 0;
  virtual void SaveSize(size_t value) = 0;
  virtual void SaveValues(size_t count, bool const *values) = 0;
  virtual void SaveValues(size_t count, int32_t const *values) = 0;
  virtual void SaveValues(size_t count, int64_t const *values) = 0;
  virtual void SaveValues(size_t count, float const *values) = 0;
  virtual void SaveValues(size_t count, double const *values) = 0;
};

class ArchiveReader {
 public:
  virtual ~ArchiveReader() = default;
  virtual std::string LoadString() = 0;
  virtual size_t LoadSize() = 0;
  virtual void LoadValues(size_t count, bool *values) = 0;
  virtual void LoadValues(size_t count, int32_t *values) = 0;
  virtual void LoadValues(size_t count, int64_t *values) = 0;
  virtual void LoadValues(size_t count, float *values) = 0;
  virtual void LoadValues(size_t count, double *values) = 0;
};

// The reference's native binary format (archive.cpp:4-83): a size is the
// 8 bytes of a size_t, a string its size then its bytes, values their
// native bytes one after the other; the machine's byte order throughout.
// A read past the end of the stream throws std::runtime_error.
class NativeBinaryArchiveWriter final : public ArchiveWriter {
 public:
  explicit NativeBinaryArchiveWriter(std::ostream &ostream) : os_(&ostream) {}
  void SaveSize(size_t value) override;
  void SaveString(std::string_view value) override;
  void SaveValues(size_t count, bool const *values) override;
  void SaveValues(size_t count, int32_t const *values) override;
  void SaveValues(size_t count, int64_t const *values) override;
  void SaveValues(size_t count, float const *values) override;
  void SaveValues(size_t count, double const *values) override;

 private:
  void Bytes(const void *p, size_t n);
  std::ostream *os_;
};

class NativeBinaryArchiveReader final : public ArchiveReader {
 public:
  explicit NativeBinaryArchiveReader(std::istream &istream) : is_(&istream) {}
  size_t LoadSize() override;
  std::string LoadString() override;
  void LoadValues(size_t count, bool *values) override;
  void LoadValues(size_t count, int32_t *values) override;
  void LoadValues(size_t count, int64_t *values) override;
  void LoadValues(size_t count, float *values) override;
  void LoadValues(size_t count, double *values) override;

 private:
  void Bytes(void *p, size_t n);
  std::istream *is_;
};

// ---- field types (component_type.cpp:8-38, shape.cpp, tensor_type.cpp) ---
enum class CType { kInvalid, kBoolean, kSInt32, kSInt64, kFloat32, kFloat64 };

template <class T>
constexpr CType CTypeOf() {
  if constexpr (std::is_same_v<T, bool>)
    return CType::kBoolean;
  else if constexpr (std::is_same_v<T, int32_t>)
    return CType::kSInt32;
  else if constexpr (std::is_same_v<T, int64_t>)
    return CType::kSInt64;
  else if constexpr (std::is_same_v<T, float>)
    return CType::kFloat32;
  else {
    static_assert(std::is_same_v<T, double>, "fields hold b, s32, s64, f32 or f64 components");
    return CType::kFloat64;
  }
}

// "b", "s32", "s64", "f32", "f64" ("INVALID" otherwise)
std::string_view CTypeName(CType c);
// inverse of CTypeName; kInvalid for any other string
CType CTypeFromString(std::string_view s);
size_t CTypeBytes(CType c);

// component type and item shape; a scalar has rank 0, a 3-vector shape {3},
// a 3x3 matrix {3, 3} (items stored column-major, as Eigen)
struct TensorType {
  CType ctype = CType::kInvalid;
  std::vector<size_t> shape;
  size_t Extent() const;
  bool operator==(const TensorType &o) const { return ctype == o.ctype && shape == o.shape; }
  bool operator!=(const TensorType &o) const { return !(*this == o); }
  void Save(ArchiveWriter &archive) const;
  void Load(ArchiveReader &archive);
};

template <class T, size_t... N>
TensorType TensorTypeOf() {
  return TensorType{CTypeOf<T>(), {N...}};
}

// ---- field storage -------------------------------------------------------
class FieldStore {
 public:
  explicit FieldStore(TensorType type)
      : type_(std::move(type)), elem_(CTypeBytes(type_.ctype)), extent_(type_.Extent()) {}
  const TensorType &GetType() const { return type_; }
  bool Same(const TensorType &t) const { return type_ == t; }
  // component type and number of components, whatever the shape
  bool Is(CType c, size_t extent) const { return type_.ctype == c && extent_ == extent; }
  size_t Extent() const { return extent_; }
  size_t ElementBytes() const { return elem_; }
  void *Data() { return bytes_.data(); }
  const void *Data() const { return bytes_.data(); }
  size_t Items() const { return extent_ ? bytes_.size() / (elem_ * extent_) : 0; }
  // new items are zero-filled (varying_field.hpp:165-167)
  void Resize(size_t n) { bytes_.resize(n * elem_ * extent_, 0); }
  // gather: new item i = old item perm[i] (varying_manager.hpp:86-102)
  void Permute(const size_t *perm, size_t n);
  // item count then the components of every item (varying_field.hpp:224-250)
  void SaveVarying(ArchiveWriter &archive) const;
  void LoadVarying(ArchiveReader &archive);
  // the components of the one item (uniform_field.hpp Save/Load)
  void SaveUniform(ArchiveWriter &archive) const;
  void LoadUniform(ArchiveReader &archive);

 private:
  void SaveComponents(ArchiveWriter &archive, size_t count) const;
  void LoadComponents(ArchiveReader &archive, size_t count);
  TensorType type_;
  size_t elem_, extent_;
  std::vector<unsigned char> bytes_;
};

template <class T, size_t... N>
class VaryingFieldSpan {
 public:
  static constexpr size_t kExtent = (size_t(1) * ... * N);
  VaryingFieldSpan() = default;
  explicit VaryingFieldSpan(FieldStore *s) : s_(s) {}
  size_t GetSize() const { return s_ ? s_->Items() : 0; }
  T *data() const { return static_cast<T *>(s_->Data()); }
  // the kExtent components of item i (column-major for matrices, as Eigen)
  T *operator[](size_t i) const { return data() + i * kExtent; }

 private:
  FieldStore *s_ = nullptr;
};

template <class T, size_t... N>
class UniformFieldSpan {
 public:
  static constexpr size_t kExtent = (size_t(1) * ... * N);
  UniformFieldSpan() = default;
  explicit UniformFieldSpan(FieldStore *s) : s_(s) {}
  T *data() const { return static_cast<T *>(s_->Data()); }
  T &operator[](size_t k) const { return data()[k]; }
  T &operator*() const { return data()[0]; }

 private:
  FieldStore *s_ = nullptr;
};

class UniformManager {
 public:
  template <class T, size_t... N>
  UniformFieldSpan<T, N...> AddFieldImpl(std::string_view name) {
    return UniformFieldSpan<T, N...>(Add(name, TensorTypeOf<T, N...>()));
  }
  template <class T, size_t... N>
  UniformFieldSpan<T, N...> FieldSpan(std::string_view name) const {
    FieldStore *f = Find(name);
    if (!f || !f->Same(TensorTypeOf<T, N...>()))
      throw std::out_of_range("field does not exist");
    return UniformFieldSpan<T, N...>(f);
  }
  // adds a field of a type known at run time (loading)
  FieldStore *AddField(std::string_view name, const TensorType &type) { return Add(name, type); }
  size_t GetFieldCount() const { return fields_.size(); }
  bool HasField(std::string_view name) const { return fields_.count(std::string(name)) != 0; }
  FieldStore *Find(std::string_view name) const;
  void RemoveField(std::string_view name) { fields_.erase(std::string(name)); }
  std::vector<std::string> GetFieldNames() const;
  // field count, then per field (ascending names) its name, type and value
  void Save(ArchiveWriter &archive) const;
  void Load(ArchiveReader &archive);

 private:
  FieldStore *Add(std::string_view name, const TensorType &type);
  std::map<std::string, std::unique_ptr<FieldStore>> fields_;
};

class VaryingManager {
 public:
  template <class T, size_t... N>
  VaryingFieldSpan<T, N...> AddFieldImpl(std::string_view name) {
    return VaryingFieldSpan<T, N...>(Add(name, TensorTypeOf<T, N...>()));
  }
  template <class T, size_t... N>
  VaryingFieldSpan<T, N...> FieldSpan(std::string_view name) const {
    FieldStore *f = Find(name);
    if (!f || !f->Same(TensorTypeOf<T, N...>()))
      throw std::out_of_range("field does not exist");
    return VaryingFieldSpan<T, N...>(f);
  }
  FieldStore *AddField(std::string_view name, const TensorType &type) { return Add(name, type); }
  size_t GetFieldCount() const { return fields_.size(); }
  bool HasField(std::string_view name) const { return fields_.count(std::string(name)) != 0; }
  FieldStore *Find(std::string_view name) const;
  void RemoveField(std::string_view name) { fields_.erase(std::string(name)); }
  size_t GetItemCount() const { return count_; }
  void ResizeItems(size_t n);
  void PermuteItems(const size_t *perm, size_t n);
  bool IsDirty() const { return dirty_; }
  void SetDirty(bool v) { dirty_ = v; }
  std::vector<std::string> GetFieldNames() const;
  // item count, field count, then per field (ascending names) its name,
  // type and values; loading sets the item count, adds the stored fields
  // (or checks the type of existing ones) and resizes each stored field to
  // its stored count (fields not in the archive keep their items, as the
  // reference's VaryingManager::Load)
  void Save(ArchiveWriter &archive) const;
  void Load(ArchiveReader &archive);

 private:
  FieldStore *Add(std::string_view name, const TensorType &type);
  std::map<std::string, std::unique_ptr<FieldStore>> fields_;
  size_t count_ = 0;
  bool dirty_ = true;
};

class Model;

class Group {
 public:
  Group(Model &model, std::string_view name, std::string_view type)
      : model_(&model), name_(name), type_(type) {}
  const Model &GetModel() const { return *model_; }
  std::string_view GetGroupName() const { return name_; }
  std::string_view GetGroupType() const { return type_; }
  GroupIndex GetGroupIndex() const { return index_; }
  void SetGroupIndex(GroupIndex v) { index_ = v; }
  size_t GetItemCount() const { return varying_.GetItemCount(); }
  bool HasTag(const std::string &tag) const { return tags_.count(tag) != 0; }
  void AddTag(const std::string &tag) { tags_.insert(tag); }
  void RemoveTag(const std::string &tag) { tags_.erase(tag); }
  std::vector<std::string> GetTags() const { return {tags_.begin(), tags_.end()}; }

  const UniformManager &GetUniform() const { return uniform_; }
  const VaryingManager &GetVarying() const { return varying_; }
  template <class T, size_t... N>
  UniformFieldSpan<T, N...> AddUniformFieldImpl(std::string_view name) {
    if (varying_.HasField(name))
      throw FieldOfDifferentKindAlreadyExistsError{};
    return uniform_.AddFieldImpl<T, N...>(name);
  }
  template <class T, size_t... N>
  VaryingFieldSpan<T, N...> AddVaryingFieldImpl(std::string_view name) {
    if (uniform_.HasField(name))
      throw FieldOfDifferentKindAlreadyExistsError{};
    return varying_.AddFieldImpl<T, N...>(name);
  }
  void RemoveField(std::string_view name) {
    uniform_.RemoveField(name);
    varying_.RemoveField(name);
  }
  void Resize(size_t n) { varying_.ResizeItems(n); }
  void Permute(const size_t *perm, size_t n) { varying_.PermuteItems(perm, n); }
  bool IsDirty() const { return varying_.IsDirty(); }
  void SetDirty(bool v) { varying_.SetDirty(v); }
  // tags, varying fields, uniform fields (name and type are the model's)
  void Save(ArchiveWriter &archive) const;
  void Load(ArchiveReader &archive);

 private:
  Model *model_;
  std::string name_, type_;
  GroupIndex index_ = GroupIndex::kNoIndex;
  std::set<std::string> tags_;
  VaryingManager varying_;
  UniformManager uniform_;
};

class Model {
 public:
  Model() = default;
  Model(const Model &) = delete;
  Model &operator=(const Model &) = delete;
  ~Model();

  // groups are indexed by the rank of their name (model.cpp:12-33)
  Group &AddGroup(std::string_view name, std::string_view type);
  Group *TryGetGroup(std::string_view name);
  void RemoveGroup(std::string_view name);
  size_t GetGroupCount() const { return groups_.size(); }
  std::vector<Group *> GetGroups();
  std::vector<std::string> GetGroupNames() const;
  Group *GetGroupByIndex(size_t index) { return by_index_.at(index); }

  bool IsDirty() const;
  void SetDirty(bool v);

  const UniformManager &GetGlobal() const { return global_; }
  template <class T, size_t... N>
  UniformFieldSpan<T, N...> AddGlobalFieldImpl(std::string_view name) {
    return global_.AddFieldImpl<T, N...>(name);
  }
  void RemoveGlobalField(std::string_view name) { global_.RemoveField(name); }

  // global fields, then per group (ascending names) its name, type and
  // contents; loading adds the stored groups (or reuses a group of the same
  // name and type) and loads into them
  void Save(ArchiveWriter &archive) const;
  void Load(ArchiveReader &archive);

 private:
  void Reindex();
  UniformManager global_;
  std::map<std::string, std::unique_ptr<Group>, std::less<>> groups_;
  std::vector<Group *> by_index_;
};

// ---- VTK output (util/save_vtk.hpp:9) --------------------------------------
// The group as ASCII legacy-VTK POLYDATA: POINTS from "position", then
// POINT_DATA velocity, acceleration (VECTORS), _index (the item index), mass,
// volume, density, pressure, time_of_birth, implicit_function (SCALARS) and
// implicit_function_gradient (VECTORS) for the fields present, reals printed
// with std::fixed. A present field that is not a float or double scalar or
// 1-, 2- or 3-vector throws std::logic_error ("not implemented", as the
// reference's NotImplementedError).
void SaveVTK(std::ostream &o, const Group &group);

// ---- neighbourhood ---------------------------------------------------------
class Neighborhood {
 public:
  Neighborhood() = default;
  void SetRadius(double radius);
  void Load(const Model &model);
  void Update();
  void Permute(Model &model);
  // neighbours of item i_ of group g_, one list per group index
  void CopyNeighbors(size_t g_, size_t i_, std::vector<std::vector<size_t>> &neighbors) const;

  double GetRadius() const { return radius_; }
  Model *GetModel() const { return model_; }

 private:
  double radius_ = 0;
  Model *model_ = nullptr;
  mutable std::map<std::pair<size_t, size_t>, std::pair<std::vector<uint32_t>, std::vector<uint32_t>>>
      csr_;
};

// ---- schemes ---------------------------------------------------------------
class SchemeBase {
 public:
  using ProcedureFunction = std::function<void(SchemeBase &, const Neighborhood &)>;
  virtual ~SchemeBase() = default;
  virtual std::string GetFullName() const = 0;
  virtual void Load(Model &model) = 0;
  void RunProcedure(std::string_view name, const Neighborhood &nhood) {
    auto it = procedures_.find(std::string(name));
    if (it == procedures_.end())
      throw "procedure does not exist";
    it->second(*this, nhood);
  }
  std::vector<std::string> GetProcedureNames() const {
    std::vector<std::string> v;
    for (const auto &kv : procedures_)
      v.push_back(kv.first);
    return v;
  }
  virtual std::string_view GetPrtclSourceCode() const { return {}; }

 protected:
  template <typename Impl>
  void RegisterProcedure(std::string_view name, void (Impl::*proc)(const Neighborhood &)) {
    procedures_.emplace(std::string(name), [proc](SchemeBase &self, const Neighborhood &nh) {
      (static_cast<Impl *>(&self)->*proc)(nh);
    });
  }
  void RegisterProcedureFunction(std::string_view name, ProcedureFunction fn) {
    procedures_.emplace(std::string(name), std::move(fn));
  }

 private:
  std::map<std::string, ProcedureFunction> procedures_;
};

class SchemeRegistry {
 public:
  SchemeRegistry(const SchemeRegistry &) = delete;
  SchemeRegistry &operator=(const SchemeRegistry &) = delete;
  template <typename Scheme>
  bool RegisterScheme(std::string_view name) {
    schemes_.emplace(std::string(name), [] { return static_cast<SchemeBase *>(new Scheme); });
    return true;
  }
  bool RegisterFactory(std::string_view name, std::function<SchemeBase *()> make) {
    schemes_.emplace(std::string(name), std::move(make));
    return true;
  }
  std::unique_ptr<SchemeBase> NewScheme(std::string_view name) {
    auto it = schemes_.find(std::string(name));
    if (it == schemes_.end())
      throw "scheme is not registered";
    return std::unique_ptr<SchemeBase>{it->second()};
  }
  std::vector<std::string> GetSchemeNames() const {
    std::vector<std::string> v;
    for (const auto &kv : schemes_)
      v.push_back(kv.first);
    return v;
  }
  friend SchemeRegistry &GetSchemeRegistry();

 private:
  SchemeRegistry() = default;
  std::map<std::string, std::function<SchemeBase *()>> schemes_;
};

SchemeRegistry &GetSchemeRegistry();

template <typename Scheme>
class SchemeRegistration {
 public:
  explicit SchemeRegistration(std::string_view name) {
    GetSchemeRegistry().RegisterScheme<Scheme>(name);
  }
};

namespace hip {

// Registers the GPU schemes (iisph, wkbb18, density, gravity, viscosity,
// symplectic_euler, aiast12) under their ", B=hip" names. Idempotent.
void RegisterSchemes();

// The device mirror of a model: one C-ABI scene for its fluid and boundary
// groups on a process-wide context (device 0).
prtcl_hip_ctx *Context();
int32_t SceneOf(Model &model);
// host spans -> device / device -> host for every mirrored field present
void Upload(Model &model);
void Download(Model &model);
// drops the device mirror of a model (also done by ~Model)
void Forget(const Model *model);

}  // namespace hip

}  // namespace prtcl
