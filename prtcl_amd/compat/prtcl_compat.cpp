// Implementation of the source-compatible host interface (prtcl_compat.hpp)
// and of the GPU schemes over the C ABI.
#include "prtcl_compat.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace prtcl {

// ---- archives ----------------------------------------------------------------
void NativeBinaryArchiveWriter::Bytes(const void *p, size_t n) {
  os_->write(static_cast<const char *>(p), static_cast<std::streamsize>(n));
  if (!*os_)
    throw std::runtime_error("archive: write failed");
}
void NativeBinaryArchiveWriter::SaveSize(size_t value) { Bytes(&value, sizeof(value)); }
void NativeBinaryArchiveWriter::SaveString(std::string_view value) {
  SaveSize(value.size());
  Bytes(value.data(), value.size());
}
void NativeBinaryArchiveWriter::SaveValues(size_t n, bool const *v) { Bytes(v, n * sizeof(*v)); }
void NativeBinaryArchiveWriter::SaveValues(size_t n, int32_t const *v) { Bytes(v, n * sizeof(*v)); }
void NativeBinaryArchiveWriter::SaveValues(size_t n, int64_t const *v) { Bytes(v, n * sizeof(*v)); }
void NativeBinaryArchiveWriter::SaveValues(size_t n, float const *v) { Bytes(v, n * sizeof(*v)); }
void NativeBinaryArchiveWriter::SaveValues(size_t n, double const *v) { Bytes(v, n * sizeof(*v)); }

void NativeBinaryArchiveReader::Bytes(void *p, size_t n) {
  is_->read(static_cast<char *>(p), static_cast<std::streamsize>(n));
  if (static_cast<size_t>(is_->gcount()) != n)
    throw std::runtime_error("archive: unexpected end of data");
}
size_t NativeBinaryArchiveReader::LoadSize() {
  size_t v = 0;
  Bytes(&v, sizeof(v));
  return v;
}
std::string NativeBinaryArchiveReader::LoadString() {
  const size_t n = LoadSize();
  std::string out(n, '\0');
  Bytes(out.data(), n);
  return out;
}
void NativeBinaryArchiveReader::LoadValues(size_t n, bool *v) { Bytes(v, n * sizeof(*v)); }
void NativeBinaryArchiveReader::LoadValues(size_t n, int32_t *v) { Bytes(v, n * sizeof(*v)); }
void NativeBinaryArchiveReader::LoadValues(size_t n, int64_t *v) { Bytes(v, n * sizeof(*v)); }
void NativeBinaryArchiveReader::LoadValues(size_t n, float *v) { Bytes(v, n * sizeof(*v)); }
void NativeBinaryArchiveReader::LoadValues(size_t n, double *v) { Bytes(v, n * sizeof(*v)); }

// ---- field types ---------------------------------------------------------------
std::string_view CTypeName(CType c) {
  switch (c) {
    case CType::kBoolean: return "b";
    case CType::kSInt32: return "s32";
    case CType::kSInt64: return "s64";
    case CType::kFloat32: return "f32";
    case CType::kFloat64: return "f64";
    default: return "INVALID";
  }
}

CType CTypeFromString(std::string_view s) {
  if (s == "b")
    return CType::kBoolean;
  if (s == "s32")
    return CType::kSInt32;
  if (s == "s64")
    return CType::kSInt64;
  if (s == "f32")
    return CType::kFloat32;
  if (s == "f64")
    return CType::kFloat64;
  return CType::kInvalid;
}

size_t CTypeBytes(CType c) {
  switch (c) {
    case CType::kBoolean: return sizeof(bool);
    case CType::kSInt32: return 4;
    case CType::kSInt64: return 8;
    case CType::kFloat32: return 4;
    case CType::kFloat64: return 8;
    default: return 0;
  }
}

size_t TensorType::Extent() const {
  size_t n = 1;
  for (size_t e : shape)
    n *= e;
  return n;
}

// component type name, then rank and extents (tensor_type.cpp:7-15,
// shape.cpp:9-20)
void TensorType::Save(ArchiveWriter &archive) const {
  archive.SaveString(CTypeName(ctype));
  archive.SaveSize(shape.size());
  for (size_t e : shape)
    archive.SaveSize(e);
}

void TensorType::Load(ArchiveReader &archive) {
  ctype = CTypeFromString(archive.LoadString());
  if (ctype == CType::kInvalid)
    throw std::runtime_error("archive: invalid component type");
  shape.resize(archive.LoadSize());
  for (size_t &e : shape)
    e = archive.LoadSize();
}

// ---- fields ----------------------------------------------------------------
void FieldStore::SaveComponents(ArchiveWriter &archive, size_t count) const {
  const void *d = bytes_.data();
  switch (type_.ctype) {
    case CType::kBoolean: archive.SaveValues(count, static_cast<const bool *>(d)); break;
    case CType::kSInt32: archive.SaveValues(count, static_cast<const int32_t *>(d)); break;
    case CType::kSInt64: archive.SaveValues(count, static_cast<const int64_t *>(d)); break;
    case CType::kFloat32: archive.SaveValues(count, static_cast<const float *>(d)); break;
    case CType::kFloat64: archive.SaveValues(count, static_cast<const double *>(d)); break;
    default: throw std::logic_error("field of invalid type");
  }
}

void FieldStore::LoadComponents(ArchiveReader &archive, size_t count) {
  void *d = bytes_.data();
  switch (type_.ctype) {
    case CType::kBoolean: archive.LoadValues(count, static_cast<bool *>(d)); break;
    case CType::kSInt32: archive.LoadValues(count, static_cast<int32_t *>(d)); break;
    case CType::kSInt64: archive.LoadValues(count, static_cast<int64_t *>(d)); break;
    case CType::kFloat32: archive.LoadValues(count, static_cast<float *>(d)); break;
    case CType::kFloat64: archive.LoadValues(count, static_cast<double *>(d)); break;
    default: throw std::logic_error("field of invalid type");
  }
}

void FieldStore::SaveVarying(ArchiveWriter &archive) const {
  const size_t n = Items();
  archive.SaveSize(n);
  if (n > 0)
    SaveComponents(archive, n * extent_);
}

void FieldStore::LoadVarying(ArchiveReader &archive) {
  const size_t n = archive.LoadSize();
  Resize(n);
  if (n > 0)
    LoadComponents(archive, n * extent_);
}

void FieldStore::SaveUniform(ArchiveWriter &archive) const { SaveComponents(archive, extent_); }

void FieldStore::LoadUniform(ArchiveReader &archive) {
  Resize(1);
  LoadComponents(archive, extent_);
}

void FieldStore::Permute(const size_t *perm, size_t n) {
  const size_t item = elem_ * extent_;
  if (n != Items())
    throw std::invalid_argument("permutation size differs from the item count");
  std::vector<unsigned char> out(bytes_.size());
  for (size_t i = 0; i < n; ++i) {
    if (perm[i] >= n)
      throw std::out_of_range("permutation index out of range");
    std::memcpy(out.data() + i * item, bytes_.data() + perm[i] * item, item);
  }
  bytes_.swap(out);
}

FieldStore *UniformManager::Find(std::string_view name) const {
  auto it = fields_.find(std::string(name));
  return it == fields_.end() ? nullptr : it->second.get();
}

FieldStore *UniformManager::Add(std::string_view name, const TensorType &type) {
  if (type.ctype == CType::kInvalid)
    throw std::invalid_argument("field of invalid type");
  auto it = fields_.find(std::string(name));
  if (it != fields_.end()) {
    if (!it->second->Same(type))
      throw FieldOfDifferentTypeAlreadyExistsError{};
    return it->second.get();
  }
  auto &slot = fields_[std::string(name)];
  slot = std::make_unique<FieldStore>(type);
  slot->Resize(1);
  return slot.get();
}

void UniformManager::Save(ArchiveWriter &archive) const {
  archive.SaveSize(fields_.size());
  for (const auto &[name, field] : fields_) {
    archive.SaveString(name);
    field->GetType().Save(archive);
    field->SaveUniform(archive);
  }
}

void UniformManager::Load(ArchiveReader &archive) {
  const size_t n = archive.LoadSize();
  for (size_t k = 0; k < n; ++k) {
    const std::string name = archive.LoadString();
    TensorType type;
    type.Load(archive);
    Add(name, type)->LoadUniform(archive);
  }
}

std::vector<std::string> UniformManager::GetFieldNames() const {
  std::vector<std::string> v;
  for (const auto &kv : fields_)
    v.push_back(kv.first);
  return v;
}

FieldStore *VaryingManager::Find(std::string_view name) const {
  auto it = fields_.find(std::string(name));
  return it == fields_.end() ? nullptr : it->second.get();
}

FieldStore *VaryingManager::Add(std::string_view name, const TensorType &type) {
  if (type.ctype == CType::kInvalid)
    throw std::invalid_argument("field of invalid type");
  auto it = fields_.find(std::string(name));
  if (it != fields_.end()) {
    if (!it->second->Same(type))
      throw FieldOfDifferentTypeAlreadyExistsError{};
    return it->second.get();
  }
  auto &slot = fields_[std::string(name)];
  slot = std::make_unique<FieldStore>(type);
  slot->Resize(count_);
  return slot.get();
}

void VaryingManager::Save(ArchiveWriter &archive) const {
  archive.SaveSize(count_);
  archive.SaveSize(fields_.size());
  for (const auto &[name, field] : fields_) {
    archive.SaveString(name);
    field->GetType().Save(archive);
    field->SaveVarying(archive);
  }
}

void VaryingManager::Load(ArchiveReader &archive) {
  count_ = archive.LoadSize();
  dirty_ = true;
  const size_t n = archive.LoadSize();
  for (size_t k = 0; k < n; ++k) {
    const std::string name = archive.LoadString();
    TensorType type;
    type.Load(archive);
    Add(name, type)->LoadVarying(archive);
  }
}

void VaryingManager::ResizeItems(size_t n) {
  for (auto &kv : fields_)
    kv.second->Resize(n);
  count_ = n;
  dirty_ = true;
}

void VaryingManager::PermuteItems(const size_t *perm, size_t n) {
  if (n != count_)
    throw std::invalid_argument("permutation size differs from the item count");
  for (auto &kv : fields_)
    kv.second->Permute(perm, n);
  dirty_ = true;
}

std::vector<std::string> VaryingManager::GetFieldNames() const {
  std::vector<std::string> v;
  for (const auto &kv : fields_)
    v.push_back(kv.first);
  return v;
}

// ---- groups and model ------------------------------------------------------
void Group::Save(ArchiveWriter &archive) const {
  archive.SaveSize(tags_.size());
  for (const std::string &tag : tags_)
    archive.SaveString(tag);
  varying_.Save(archive);
  uniform_.Save(archive);
}

void Group::Load(ArchiveReader &archive) {
  const size_t n = archive.LoadSize();
  for (size_t k = 0; k < n; ++k)
    tags_.insert(archive.LoadString());
  varying_.Load(archive);
  uniform_.Load(archive);
}

Model::~Model() { hip::Forget(this); }

void Model::Save(ArchiveWriter &archive) const {
  global_.Save(archive);
  archive.SaveSize(groups_.size());
  for (const auto &[name, group] : groups_) {
    archive.SaveString(name);
    archive.SaveString(group->GetGroupType());
    group->Save(archive);
  }
}

void Model::Load(ArchiveReader &archive) {
  global_.Load(archive);
  const size_t n = archive.LoadSize();
  for (size_t k = 0; k < n; ++k) {
    const std::string name = archive.LoadString();
    const std::string type = archive.LoadString();
    AddGroup(name, type).Load(archive);
  }
}

Group &Model::AddGroup(std::string_view name, std::string_view type) {
  if (name.empty())
    throw std::invalid_argument("invalid identifier");
  auto it = groups_.find(name);
  if (it == groups_.end()) {
    it = groups_.emplace(std::string(name), std::make_unique<Group>(*this, name, type)).first;
  } else if (it->second->GetGroupType() != type) {
    throw GroupOfDifferentTypeAlreadyExists{};
  }
  Reindex();
  return *it->second;
}

void Model::Reindex() {
  by_index_.clear();
  for (auto &kv : groups_) {
    kv.second->SetGroupIndex(static_cast<GroupIndex>(by_index_.size()));
    by_index_.push_back(kv.second.get());
  }
}

Group *Model::TryGetGroup(std::string_view name) {
  auto it = groups_.find(name);
  return it == groups_.end() ? nullptr : it->second.get();
}

void Model::RemoveGroup(std::string_view name) {
  auto it = groups_.find(name);
  if (it != groups_.end()) {
    groups_.erase(it);
    Reindex();
  }
}

std::vector<Group *> Model::GetGroups() { return by_index_; }

std::vector<std::string> Model::GetGroupNames() const {
  std::vector<std::string> v;
  for (const auto &kv : groups_)
    v.push_back(kv.first);
  return v;
}

bool Model::IsDirty() const {
  for (const auto &kv : groups_)
    if (kv.second->IsDirty())
      return true;
  return false;
}

void Model::SetDirty(bool v) {
  for (auto &kv : groups_)
    kv.second->SetDirty(v);
}

// ---- VTK output ----------------------------------------------------------------
namespace {

// the items of a float / double field of 1, 2 or 3 components, one line
// each, components separated by a space (save_vtk.cpp:11-53)
template <class T>
void vtk_lines(std::ostream &o, const FieldStore &f) {
  const T *d = static_cast<const T *>(f.Data());
  const size_t nc = f.Extent();
  for (size_t i = 0; i < f.Items(); ++i) {
    for (size_t c = 0; c < nc; ++c) {
      if (c)
        o << ' ';
      o << std::fixed << d[i * nc + c];
    }
    o << '\n';
  }
}

void vtk_field(std::ostream &o, const VaryingManager &v, const char *name) {
  const FieldStore *f = v.Find(name);
  if (!f)
    throw std::out_of_range("field does not exist");
  const TensorType &t = f->GetType();
  const bool shape_ok = t.shape.empty() || (t.shape.size() == 1 && t.shape[0] >= 1 &&
                                            t.shape[0] <= 3);
  if (shape_ok && t.ctype == CType::kFloat32)
    return vtk_lines<float>(o, *f);
  if (shape_ok && t.ctype == CType::kFloat64)
    return vtk_lines<double>(o, *f);
  throw std::logic_error("not implemented");
}

}  // namespace

void SaveVTK(std::ostream &o, const Group &group) {
  const VaryingManager &v = group.GetVarying();
  const size_t n = group.GetItemCount();
  o << "# vtk DataFile Version 2.0\n";
  o << group.GetGroupName() << " " << group.GetGroupType() << "\n";
  o << "ASCII\n";
  o << "DATASET POLYDATA\n";
  const auto flags = o.flags();
  o << "POINTS " << n << " float\n";
  vtk_field(o, v, "position");
  o << "POINT_DATA " << n << "\n";
  auto vectors = [&](const char *name) {
    if (v.HasField(name)) {
      o << "VECTORS " << name << " float\n";
      vtk_field(o, v, name);
    }
  };
  auto scalars = [&](const char *name) {
    if (v.HasField(name)) {
      o << "SCALARS " << name << " float 1\n";
      o << "LOOKUP_TABLE default\n";
      vtk_field(o, v, name);
    }
  };
  vectors("velocity");
  vectors("acceleration");
  o << "SCALARS _index float 1\n";
  o << "LOOKUP_TABLE default\n";
  for (size_t i = 0; i < n; ++i)
    o << std::fixed << i << "\n";
  for (const char *name : {"mass", "volume", "density", "pressure", "time_of_birth",
                           "implicit_function"})
    scalars(name);
  vectors("implicit_function_gradient");
  o.flags(flags);
}

// ---- schemes registry --------------------------------------------------------
SchemeRegistry &GetSchemeRegistry() {
  static SchemeRegistry registry;
  return registry;
}

// ---- device mirror -----------------------------------------------------------
namespace hip {
namespace {

void check(prtcl_hip_ctx *ctx, int rc, const char *what) {
  if (rc != PRTCL_HIP_OK && rc != PRTCL_HIP_NOT_CONVERGED)
    throw std::runtime_error(std::string(what) + ": " + prtcl_hip_last_error(ctx));
}

struct Mirror {
  int32_t scene = -1;
  prtcl_hip_scene_desc desc{};
  double radius = 0;  // from the Neighborhood; 0: 2h
};

std::mutex g_mu;
std::map<const Model *, Mirror> g_mirrors;

struct Mapped {
  int group;
  int field;
  const char *name;
  size_t extent;
};
const Mapped kMapped[] = {
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_POSITION, "position", 3},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_VELOCITY, "velocity", 3},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_ACCELERATION, "acceleration", 3},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_DENSITY, "density", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_PRESSURE, "pressure", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_MASS, "mass", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_IISPH_HELPER_C, "iisph_helper_c", 3},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_IISPH_SOURCE_TERM, "iisph_source_term", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_IISPH_DIAGONAL_ELEMENT, "iisph_diagonal_element", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_IISPH_RIGHT_HAND_SIDE, "iisph_right_hand_side", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_TIME_OF_BIRTH, "time_of_birth", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_WKBB18_VELOCITY_DELTA, "wkbb18_velocity_delta", 3},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_WKBB18_SYSTEM_DIAGONAL, "wkbb18_system_diagonal", 9},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_GRADIENT_CORRECTION, "gradient_correction", 9},
    // pt16 (a scene mirrors them when the model has pt16's fields)
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_PT16_NORMALIZER, "pt16_normalizer", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_PT16_DENSITY, "pt16_density", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_PT16_SAME_DENSITY, "pt16_same_density", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_VORTICITY, "vorticity", 3},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_TARGET_VELOCITY_GRADIENT, "target_velocity_gradient", 9},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_PT16_VORTICITY_DIFFUSION_DIAGONAL,
     "pt16_vorticity_diffusion_diagonal", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_PT16_VORTICITY_DIFFUSION_RHS, "pt16_vorticity_diffusion_rhs",
     3},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_PT16_VELOCITY_RECONSTRUCTION_DIAGONAL,
     "pt16_velocity_reconstruction_diagonal", 1},
    {PRTCL_HIP_FLUID, PRTCL_HIP_F_PT16_VELOCITY_RECONSTRUCTION_RHS,
     "pt16_velocity_reconstruction_rhs", 3},
    {PRTCL_HIP_BOUNDARY, PRTCL_HIP_F_POSITION, "position", 3},
    {PRTCL_HIP_BOUNDARY, PRTCL_HIP_F_VOLUME, "volume", 1},
};

Group *group_of_type(Model &m, std::string_view type) {
  Group *found = nullptr;
  for (Group *g : m.GetGroups())
    if (g->GetGroupType() == type) {
      if (found)
        throw std::invalid_argument("the device mirror supports one group of each type");
      found = g;
    }
  return found;
}

template <class T>
bool read_uniform(const UniformManager &u, const char *name, T *out, size_t k = 0) {
  FieldStore *f = u.Find(name);
  if (!f)
    return false;
  if (k >= f->Extent())
    return false;
  const CType c = f->GetType().ctype;
  if (c == CType::kFloat32)
    *out = static_cast<T>(static_cast<const float *>(f->Data())[k]);
  else if (c == CType::kFloat64)
    *out = static_cast<T>(static_cast<const double *>(f->Data())[k]);
  else if (c == CType::kSInt32)
    *out = static_cast<T>(static_cast<const int32_t *>(f->Data())[k]);
  else
    return false;
  return true;
}

template <class T>
void write_uniform(const UniformManager &u, const char *name, T v) {
  FieldStore *f = u.Find(name);
  if (!f || f->Extent() != 1)
    return;
  const CType c = f->GetType().ctype;
  if (c == CType::kFloat32)
    *static_cast<float *>(f->Data()) = static_cast<float>(v);
  else if (c == CType::kFloat64)
    *static_cast<double *>(f->Data()) = static_cast<double>(v);
  else if (c == CType::kSInt32)
    *static_cast<int32_t *>(f->Data()) = static_cast<int32_t>(v);
}

prtcl_hip_scene_desc desc_of(Model &m, double radius) {
  prtcl_hip_scene_desc d;
  prtcl_hip_scene_desc_init(&d);
  Group *f = group_of_type(m, "fluid");
  Group *b = group_of_type(m, "boundary");
  d.n_fluid = f ? static_cast<int64_t>(f->GetItemCount()) : 0;
  d.n_boundary = b ? static_cast<int64_t>(b->GetItemCount()) : 0;
  const UniformManager &g = m.GetGlobal();
  read_uniform(g, "smoothing_scale", &d.smoothing_scale);
  read_uniform(g, "time_step", &d.time_step);
  read_uniform(g, "iisph_relaxation", &d.iisph_relaxation);
  read_uniform(g, "fade_duration", &d.fade_duration);
  for (int k = 0; k < 3; ++k) {
    read_uniform(g, "gravity", &d.gravity[k], k);
    read_uniform(g, "gravity_center", &d.gravity_center[k], k);
  }
  if (f) {
    read_uniform(f->GetUniform(), "rest_density", &d.rest_density);
    read_uniform(f->GetUniform(), "dynamic_viscosity", &d.fluid_viscosity);
    read_uniform(f->GetUniform(), "wkbb18_maximum_error", &d.wkbb18_maximum_error);
    read_uniform(f->GetUniform(), "wkbb18_maximum_iterations", &d.wkbb18_maximum_iterations);
    // a loaded wkbb18 / wkbb18_gc scheme: the scene carries the PCG vectors
    if (f->GetVarying().HasField("wkbb18_system_diagonal"))
      d.implicit_viscosity = f->GetVarying().HasField("gradient_correction") ? 2 : 1;
    // a loaded pt16 scheme: the scene carries pt16's fields (and PCG vectors)
    if (f->GetVarying().HasField("pt16_same_density"))
      d.implicit_viscosity = 3;
  }
  if (b)
    read_uniform(b->GetUniform(), "dynamic_viscosity", &d.boundary_viscosity);
  d.radius = radius;
  d.permute_every = 0;
  d.step_mode = 1;
  return d;
}

bool same_desc(const prtcl_hip_scene_desc &a, const prtcl_hip_scene_desc &b) {
  return std::memcmp(&a, &b, sizeof(a)) == 0;
}

// the scene of a model, created or re-created when the model's shape or
// parameters changed (all mirrored fields uploaded then)
int32_t scene_locked(Model &m, Mirror &mr) {
  prtcl_hip_ctx *ctx = Context();
  const prtcl_hip_scene_desc d = desc_of(m, mr.radius);
  if (mr.scene >= 0 && same_desc(d, mr.desc))
    return mr.scene;
  if (mr.scene >= 0)
    check(ctx, prtcl_hip_scene_destroy(ctx, mr.scene), "scene_destroy");
  mr.scene = -1;
  int32_t id = -1;
  prtcl_hip_scene_desc dd = d;
  check(ctx, prtcl_hip_scene_create(ctx, &dd, &id), "scene_create");
  mr.scene = id;
  mr.desc = d;
  return id;
}

void transfer(Model &m, int32_t scene, bool up) {
  prtcl_hip_ctx *ctx = Context();
  Group *groups[2] = {group_of_type(m, "boundary"), group_of_type(m, "fluid")};
  for (const Mapped &mf : kMapped) {
    Group *g = groups[mf.group];
    if (!g || g->GetItemCount() == 0)
      continue;
    FieldStore *fs = g->GetVarying().Find(mf.name);
    if (!fs || !fs->Is(CType::kFloat32, mf.extent))
      continue;
    const size_t count = g->GetItemCount() * mf.extent;
    if (up)
      check(ctx, prtcl_hip_upload_field(ctx, scene, mf.group, mf.field, fs->Data(), count),
            "upload_field");
    else
      check(ctx, prtcl_hip_download_field(ctx, scene, mf.group, mf.field, fs->Data(), count),
            "download_field");
  }
  const UniformManager &gl = m.GetGlobal();
  if (up) {
    // globals the host loop sets (example.lua:368-369, 414)
    double v = 0;
    if (read_uniform(gl, "current_time", &v))
      check(ctx, prtcl_hip_set_global(ctx, scene, PRTCL_HIP_G_CURRENT_TIME, v), "set_global");
    if (read_uniform(gl, "iisph_aprde", &v))
      check(ctx, prtcl_hip_set_global(ctx, scene, PRTCL_HIP_G_IISPH_APRDE, v), "set_global");
    if (read_uniform(gl, "iisph_nprde", &v))
      check(ctx, prtcl_hip_set_global(ctx, scene, PRTCL_HIP_G_IISPH_NPRDE, v), "set_global");
    // pt16's fluid uniforms, set by the host loop between its solves
    // (example.lua:436-443)
    Group *f = groups[PRTCL_HIP_FLUID];
    if (f && f->GetVarying().HasField("pt16_same_density")) {
      const UniformManager &u = f->GetUniform();
      if (read_uniform(u, "pt16_maximum_error", &v))
        check(ctx, prtcl_hip_set_global(ctx, scene, PRTCL_HIP_G_PT16_MAXIMUM_ERROR, v),
              "set_global");
      if (read_uniform(u, "pt16_maximum_iterations", &v))
        check(ctx, prtcl_hip_set_global(ctx, scene, PRTCL_HIP_G_PT16_MAXIMUM_ITERATIONS, v),
              "set_global");
      if (read_uniform(u, "strain_rate_viscosity", &v))
        check(ctx, prtcl_hip_set_global(ctx, scene, PRTCL_HIP_G_STRAIN_RATE_VISCOSITY, v),
              "set_global");
    }
  } else {
    double v = 0;
    check(ctx, prtcl_hip_get_global(ctx, scene, PRTCL_HIP_G_IISPH_APRDE, &v), "get_global");
    write_uniform(gl, "iisph_aprde", v);
    check(ctx, prtcl_hip_get_global(ctx, scene, PRTCL_HIP_G_IISPH_NPRDE, &v), "get_global");
    write_uniform(gl, "iisph_nprde", v);
    check(ctx, prtcl_hip_get_global(ctx, scene, PRTCL_HIP_G_MAXIMUM_SPEED, &v), "get_global");
    write_uniform(gl, "maximum_speed", v);
    if (Group *f = groups[PRTCL_HIP_FLUID]) {
      check(ctx, prtcl_hip_get_global(ctx, scene, PRTCL_HIP_G_WKBB18_ITERATIONS, &v),
            "get_global");
      write_uniform(f->GetUniform(), "wkbb18_iterations", v);
      if (f->GetVarying().HasField("pt16_same_density")) {
        check(ctx, prtcl_hip_get_global(ctx, scene, PRTCL_HIP_G_PT16_ITERATIONS, &v),
              "get_global");
        write_uniform(f->GetUniform(), "pt16_iterations", v);
      }
    }
  }
}

int device_group(Model &m, size_t group_index) {
  Group *g = m.GetGroupByIndex(group_index);
  if (g->GetGroupType() == "fluid")
    return PRTCL_HIP_FLUID;
  if (g->GetGroupType() == "boundary")
    return PRTCL_HIP_BOUNDARY;
  return -1;
}

}  // namespace

prtcl_hip_ctx *Context() {
  static prtcl_hip_ctx *ctx = [] {
    prtcl_hip_ctx *c = nullptr;
    if (prtcl_hip_ctx_create(0, 0, &c) != PRTCL_HIP_OK)
      throw std::runtime_error("prtcl_hip_ctx_create failed (no MI355X device?)");
    return c;
  }();
  return ctx;
}

int32_t SceneOf(Model &model) {
  std::lock_guard<std::mutex> lk(g_mu);
  Mirror &mr = g_mirrors[&model];
  const int32_t old = mr.scene;
  const int32_t id = scene_locked(model, mr);
  if (id != old)
    transfer(model, id, true);
  return id;
}

void Upload(Model &model) { transfer(model, SceneOf(model), true); }
void Download(Model &model) { transfer(model, SceneOf(model), false); }

void Forget(const Model *model) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_mirrors.find(model);
  if (it == g_mirrors.end())
    return;
  if (it->second.scene >= 0)
    (void)prtcl_hip_scene_destroy(Context(), it->second.scene);
  g_mirrors.erase(it);
}

namespace {

void set_radius(Model &m, double r) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_mirrors[&m].radius = r;
}

// run one procedure with the host spans authoritative before and after
void run(Model &m, int proc) {
  prtcl_hip_ctx *ctx = Context();
  const int32_t sc = SceneOf(m);
  transfer(m, sc, true);
  check(ctx, prtcl_hip_run(ctx, sc, proc), "run");
  transfer(m, sc, false);
}

}  // namespace
}  // namespace hip

// ---- neighbourhood -----------------------------------------------------------
void Neighborhood::SetRadius(double radius) {
  radius_ = radius;
  if (model_)
    hip::set_radius(*model_, radius);
}

void Neighborhood::Load(const Model &model) {
  model_ = const_cast<Model *>(&model);
  hip::set_radius(*model_, radius_);
  csr_.clear();
}

void Neighborhood::Update() {
  if (!model_)
    throw std::logic_error("neighborhood has no model");
  hip::run(*model_, PRTCL_HIP_P_NSEARCH_UPDATE);
  csr_.clear();
}

void Neighborhood::Permute(Model &model) {
  if (&model != model_)
    Load(model);
  prtcl_hip_ctx *ctx = hip::Context();
  const int32_t sc = hip::SceneOf(model);
  hip::transfer(model, sc, true);
  hip::check(ctx, prtcl_hip_run(ctx, sc, PRTCL_HIP_P_NSEARCH_PERMUTE), "permute");
  // the device's host order is now its slot order: apply the same gather to
  // every varying field of the mirrored groups
  for (const char *type : {"boundary", "fluid"}) {
    Group *g = hip::group_of_type(model, type);
    if (!g || g->GetItemCount() == 0)
      continue;
    const int dg = std::string_view(type) == "fluid" ? PRTCL_HIP_FLUID : PRTCL_HIP_BOUNDARY;
    std::vector<uint32_t> p(g->GetItemCount());
    hip::check(ctx, prtcl_hip_last_permutation(ctx, sc, dg, p.data()), "last_permutation");
    std::vector<size_t> perm(p.begin(), p.end());
    g->Permute(perm.data(), perm.size());
  }
  csr_.clear();
}

void Neighborhood::CopyNeighbors(size_t g_, size_t i_,
                                 std::vector<std::vector<size_t>> &neighbors) const {
  if (!model_)
    throw std::logic_error("neighborhood has no model");
  prtcl_hip_ctx *ctx = hip::Context();
  const int32_t sc = hip::SceneOf(*model_);
  const size_t ng = model_->GetGroupCount();
  neighbors.resize(ng);
  const int qg = hip::device_group(*model_, g_);
  for (size_t h = 0; h < ng; ++h) {
    neighbors[h].clear();
    const int dg = hip::device_group(*model_, h);
    if (qg < 0 || dg < 0)
      continue;
    auto key = std::make_pair(g_, h);
    auto it = csr_.find(key);
    if (it == csr_.end()) {
      const size_t nq = model_->GetGroupByIndex(g_)->GetItemCount();
      std::vector<uint32_t> off(nq + 1);
      uint64_t total = 0;
      hip::check(ctx, prtcl_hip_copy_neighbors(ctx, sc, qg, dg, off.data(), nullptr, &total),
                 "copy_neighbors");
      std::vector<uint32_t> idx(total);
      hip::check(ctx, prtcl_hip_copy_neighbors(ctx, sc, qg, dg, off.data(), idx.data(), &total),
                 "copy_neighbors");
      it = csr_.emplace(key, std::make_pair(std::move(off), std::move(idx))).first;
    }
    const auto &off = it->second.first;
    const auto &idx = it->second.second;
    if (i_ + 1 >= off.size())
      throw std::out_of_range("item index out of range");
    neighbors[h].assign(idx.begin() + off[i_], idx.begin() + off[i_ + 1]);
  }
}

// ---- GPU schemes ---------------------------------------------------------------
namespace hip {
namespace {

using Binder = std::function<void(Model &)>;

class HipScheme final : public SchemeBase {
 public:
  HipScheme(std::string name, const std::vector<std::pair<const char *, int>> &procs,
            Binder bind)
      : name_(std::move(name)), bind_(std::move(bind)) {
    for (const auto &p : procs) {
      const int proc = p.second;
      RegisterProcedureFunction(p.first, [proc](SchemeBase &self, const Neighborhood &nh) {
        static_cast<HipScheme &>(self).Run(proc, nh);
      });
    }
  }
  std::string GetFullName() const override { return name_; }
  void Load(Model &model) override {
    bind_(model);
    model_ = &model;
  }
  void Run(int proc, const Neighborhood &nh) {
    if (!model_)
      throw std::logic_error("scheme is not loaded");
    if (nh.GetModel() && nh.GetModel() != model_)
      throw std::logic_error("neighborhood belongs to another model");
    run(*model_, proc);
  }

 private:
  std::string name_;
  Binder bind_;
  Model *model_ = nullptr;
};

template <class Fn>
void each_of_type(Model &m, std::string_view type, Fn &&fn) {
  for (Group *g : m.GetGroups())
    if (g->GetGroupType() == type)
      fn(*g);
}

template <class Fn>
void each_dynamic(Model &m, Fn &&fn) {
  for (Group *g : m.GetGroups())
    if (g->HasTag("dynamic"))
      fn(*g);
}

std::string name_of(const char *scheme) {
  return std::string("prtcl::schemes::") + scheme + "[T=f32, N=3, K=Cubic Spline, B=hip]";
}

}  // namespace

void RegisterSchemes() {
  static std::once_flag once;
  std::call_once(once, [] {
    SchemeRegistry &R = GetSchemeRegistry();
    // field bindings of the reference schemes' Load (iisph.hpp:134-185,
    // wkbb18.hpp:135-188 and the glue schemes)
    R.RegisterFactory(name_of("iisph"), [] {
      return new HipScheme(
          name_of("iisph"),
          {{"setup", PRTCL_HIP_P_IISPH_SETUP},
           {"iteration_pressure_acceleration", PRTCL_HIP_P_IISPH_ITERATION_PRESSURE_ACCELERATION},
           {"iteration_pressure", PRTCL_HIP_P_IISPH_ITERATION_PRESSURE}},
          [](Model &m) {
            m.AddGlobalFieldImpl<float, 3>("gravity");
            m.AddGlobalFieldImpl<float>("smoothing_scale");
            m.AddGlobalFieldImpl<float>("time_step");
            m.AddGlobalFieldImpl<float>("iisph_relaxation");
            m.AddGlobalFieldImpl<float>("iisph_aprde");
            m.AddGlobalFieldImpl<int32_t>("iisph_nprde");
            each_of_type(m, "fluid", [](Group &g) {
              g.AddUniformFieldImpl<float>("rest_density");
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float, 3>("velocity");
              g.AddVaryingFieldImpl<float, 3>("acceleration");
              g.AddVaryingFieldImpl<float>("density");
              g.AddVaryingFieldImpl<float>("pressure");
              g.AddVaryingFieldImpl<float>("mass");
              g.AddVaryingFieldImpl<float, 3>("iisph_helper_c");
              g.AddVaryingFieldImpl<float>("iisph_source_term");
              g.AddVaryingFieldImpl<float>("iisph_diagonal_element");
              g.AddVaryingFieldImpl<float>("iisph_right_hand_side");
            });
            each_of_type(m, "boundary", [](Group &g) {
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float>("volume");
            });
          });
    });
    R.RegisterFactory(name_of("wkbb18"), [] {
      return new HipScheme(
          name_of("wkbb18"),
          {{"compute_diagonal", PRTCL_HIP_P_WKBB18_COMPUTE_DIAGONAL},
           {"accumulate_acceleration", PRTCL_HIP_P_WKBB18_ACCUMULATE_ACCELERATION}},
          [](Model &m) {
            m.AddGlobalFieldImpl<float>("smoothing_scale");
            m.AddGlobalFieldImpl<float>("time_step");
            m.AddGlobalFieldImpl<float>("current_time");
            m.AddGlobalFieldImpl<float>("fade_duration");
            each_of_type(m, "fluid", [](Group &g) {
              g.AddUniformFieldImpl<float>("rest_density");
              g.AddUniformFieldImpl<float>("dynamic_viscosity");
              g.AddUniformFieldImpl<float>("wkbb18_maximum_error");
              g.AddUniformFieldImpl<int32_t>("wkbb18_maximum_iterations");
              g.AddUniformFieldImpl<int32_t>("wkbb18_iterations");
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float, 3>("velocity");
              g.AddVaryingFieldImpl<float, 3>("acceleration");
              g.AddVaryingFieldImpl<float>("density");
              g.AddVaryingFieldImpl<float>("mass");
              g.AddVaryingFieldImpl<float, 3>("wkbb18_velocity_delta");
              g.AddVaryingFieldImpl<float, 3, 3>("wkbb18_system_diagonal");
              g.AddVaryingFieldImpl<float>("time_of_birth");
            });
            each_of_type(m, "boundary", [](Group &g) {
              g.AddUniformFieldImpl<float>("dynamic_viscosity");
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float>("volume");
            });
          });
    });
    // wkbb18_gc.hpp:135-190: the wkbb18 bindings plus gradient_correction
    R.RegisterFactory(name_of("wkbb18_gc"), [] {
      return new HipScheme(
          name_of("wkbb18_gc"),
          {{"compute_diagonal", PRTCL_HIP_P_WKBB18_GC_COMPUTE_DIAGONAL},
           {"accumulate_acceleration", PRTCL_HIP_P_WKBB18_GC_ACCUMULATE_ACCELERATION}},
          [](Model &m) {
            m.AddGlobalFieldImpl<float>("smoothing_scale");
            m.AddGlobalFieldImpl<float>("time_step");
            m.AddGlobalFieldImpl<float>("current_time");
            m.AddGlobalFieldImpl<float>("fade_duration");
            each_of_type(m, "fluid", [](Group &g) {
              g.AddUniformFieldImpl<float>("rest_density");
              g.AddUniformFieldImpl<float>("dynamic_viscosity");
              g.AddUniformFieldImpl<float>("wkbb18_maximum_error");
              g.AddUniformFieldImpl<int32_t>("wkbb18_maximum_iterations");
              g.AddUniformFieldImpl<int32_t>("wkbb18_iterations");
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float, 3>("velocity");
              g.AddVaryingFieldImpl<float, 3>("acceleration");
              g.AddVaryingFieldImpl<float>("density");
              g.AddVaryingFieldImpl<float>("mass");
              g.AddVaryingFieldImpl<float, 3, 3>("gradient_correction");
              g.AddVaryingFieldImpl<float, 3>("wkbb18_velocity_delta");
              g.AddVaryingFieldImpl<float, 3, 3>("wkbb18_system_diagonal");
              g.AddVaryingFieldImpl<float>("time_of_birth");
            });
            each_of_type(m, "boundary", [](Group &g) {
              g.AddUniformFieldImpl<float>("dynamic_viscosity");
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float>("volume");
            });
          });
    });
    // pt16.hpp:140-198
    R.RegisterFactory(name_of("pt16"), [] {
      return new HipScheme(
          name_of("pt16"),
          {{"setup", PRTCL_HIP_P_PT16_SETUP},
           {"solve_vorticity_diffusion", PRTCL_HIP_P_PT16_SOLVE_VORTICITY_DIFFUSION},
           {"vorticity_preservation", PRTCL_HIP_P_PT16_VORTICITY_PRESERVATION},
           {"solve_velocity_reconstruction", PRTCL_HIP_P_PT16_SOLVE_VELOCITY_RECONSTRUCTION}},
          [](Model &m) {
            m.AddGlobalFieldImpl<float>("smoothing_scale");
            m.AddGlobalFieldImpl<float>("time_step");
            m.AddGlobalFieldImpl<float>("current_time");
            m.AddGlobalFieldImpl<float>("fade_duration");
            each_of_type(m, "fluid", [](Group &g) {
              g.AddUniformFieldImpl<float>("rest_density");
              g.AddUniformFieldImpl<float>("strain_rate_viscosity");
              g.AddUniformFieldImpl<float>("pt16_maximum_error");
              g.AddUniformFieldImpl<int32_t>("pt16_maximum_iterations");
              g.AddUniformFieldImpl<int32_t>("pt16_iterations");
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float, 3>("velocity");
              g.AddVaryingFieldImpl<float, 3>("acceleration");
              g.AddVaryingFieldImpl<float>("density");
              g.AddVaryingFieldImpl<float>("mass");
              g.AddVaryingFieldImpl<float, 3, 3>("gradient_correction");
              g.AddVaryingFieldImpl<float>("time_of_birth");
              g.AddVaryingFieldImpl<float>("pt16_normalizer");
              g.AddVaryingFieldImpl<float>("pt16_density");
              g.AddVaryingFieldImpl<float>("pt16_same_density");
              g.AddVaryingFieldImpl<float, 3>("vorticity");
              g.AddVaryingFieldImpl<float, 3, 3>("target_velocity_gradient");
              g.AddVaryingFieldImpl<float>("pt16_vorticity_diffusion_diagonal");
              g.AddVaryingFieldImpl<float, 3>("pt16_vorticity_diffusion_rhs");
              g.AddVaryingFieldImpl<float>("pt16_velocity_reconstruction_diagonal");
              g.AddVaryingFieldImpl<float, 3>("pt16_velocity_reconstruction_rhs");
            });
            each_of_type(m, "boundary", [](Group &g) {
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float>("volume");
            });
          });
    });
    // correction.hpp:85-130
    R.RegisterFactory(name_of("correction"), [] {
      return new HipScheme(
          name_of("correction"),
          {{"identity_gradient_correction", PRTCL_HIP_P_CORRECTION_IDENTITY_GRADIENT_CORRECTION},
           {"compute_gradient_correction", PRTCL_HIP_P_CORRECTION_COMPUTE_GRADIENT_CORRECTION}},
          [](Model &m) {
            m.AddGlobalFieldImpl<float>("smoothing_scale");
            each_of_type(m, "fluid", [](Group &g) {
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float>("density");
              g.AddVaryingFieldImpl<float>("mass");
              g.AddVaryingFieldImpl<float, 3, 3>("gradient_correction");
            });
          });
    });
    R.RegisterFactory(name_of("density"), [] {
      return new HipScheme(name_of("density"),
                           {{"compute_density", PRTCL_HIP_P_DENSITY_COMPUTE_DENSITY}},
                           [](Model &m) {
                             m.AddGlobalFieldImpl<float>("smoothing_scale");
                             each_of_type(m, "fluid", [](Group &g) {
                               g.AddUniformFieldImpl<float>("rest_density");
                               g.AddVaryingFieldImpl<float, 3>("position");
                               g.AddVaryingFieldImpl<float>("density");
                               g.AddVaryingFieldImpl<float>("mass");
                             });
                             each_of_type(m, "boundary", [](Group &g) {
                               g.AddVaryingFieldImpl<float, 3>("position");
                               g.AddVaryingFieldImpl<float>("volume");
                             });
                           });
    });
    R.RegisterFactory(name_of("gravity"), [] {
      return new HipScheme(name_of("gravity"),
                           {{"initialize_acceleration",
                             PRTCL_HIP_P_GRAVITY_INITIALIZE_ACCELERATION},
                            {"accumulate_acceleration",
                             PRTCL_HIP_P_GRAVITY_ACCUMULATE_ACCELERATION},
                            {"initialize_axial_acceleration",
                             PRTCL_HIP_P_GRAVITY_INITIALIZE_AXIAL_ACCELERATION}},
                           [](Model &m) {
                             m.AddGlobalFieldImpl<float, 3>("gravity");
                             m.AddGlobalFieldImpl<float, 3>("gravity_center");
                             each_dynamic(m, [](Group &g) {
                               g.AddVaryingFieldImpl<float, 3>("position");
                               g.AddVaryingFieldImpl<float, 3>("acceleration");
                             });
                           });
    });
    R.RegisterFactory(name_of("viscosity"), [] {
      return new HipScheme(name_of("viscosity"),
                           {{"accumulate_acceleration",
                             PRTCL_HIP_P_VISCOSITY_ACCUMULATE_ACCELERATION}},
                           [](Model &m) {
                             m.AddGlobalFieldImpl<float>("smoothing_scale");
                             each_of_type(m, "fluid", [](Group &g) {
                               g.AddUniformFieldImpl<float>("rest_density");
                               g.AddUniformFieldImpl<float>("dynamic_viscosity");
                               g.AddVaryingFieldImpl<float, 3>("position");
                               g.AddVaryingFieldImpl<float, 3>("velocity");
                               g.AddVaryingFieldImpl<float, 3>("acceleration");
                               g.AddVaryingFieldImpl<float>("density");
                               g.AddVaryingFieldImpl<float>("mass");
                             });
                             each_of_type(m, "boundary", [](Group &g) {
                               g.AddUniformFieldImpl<float>("dynamic_viscosity");
                               g.AddVaryingFieldImpl<float, 3>("position");
                               g.AddVaryingFieldImpl<float>("volume");
                             });
                           });
    });
    R.RegisterFactory(name_of("symplectic_euler"), [] {
      return new HipScheme(
          name_of("symplectic_euler"),
          {{"integrate_velocity_with_hard_fade", PRTCL_HIP_P_SE_INTEGRATE_VELOCITY_WITH_HARD_FADE},
           {"integrate_position", PRTCL_HIP_P_SE_INTEGRATE_POSITION}},
          [](Model &m) {
            m.AddGlobalFieldImpl<float>("time_step");
            m.AddGlobalFieldImpl<float>("current_time");
            m.AddGlobalFieldImpl<float>("fade_duration");
            m.AddGlobalFieldImpl<float>("maximum_speed");
            each_dynamic(m, [](Group &g) {
              g.AddVaryingFieldImpl<float, 3>("position");
              g.AddVaryingFieldImpl<float, 3>("velocity");
              g.AddVaryingFieldImpl<float, 3>("acceleration");
              g.AddVaryingFieldImpl<float>("time_of_birth");
            });
          });
    });
    R.RegisterFactory(name_of("aiast12"), [] {
      return new HipScheme(name_of("aiast12"),
                           {{"compute_volume", PRTCL_HIP_P_AIAST12_COMPUTE_VOLUME}},
                           [](Model &m) {
                             m.AddGlobalFieldImpl<float>("smoothing_scale");
                             each_of_type(m, "boundary", [](Group &g) {
                               g.AddVaryingFieldImpl<float, 3>("position");
                               g.AddVaryingFieldImpl<float>("volume");
                             });
                           });
    });
  });
}

}  // namespace hip
}  // namespace prtcl
