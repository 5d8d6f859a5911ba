// Checkpoint and resume of device-resident scenes in the reference's native
// binary model format. Host code: the fields come off the device through the
// scene's gather-to-host-order kernel into a device staging buffer, then
// into one of two page-locked buffers, so the copy of the next field
// overlaps the file write of the current one.
#include "checkpoint.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace prtcl_amd {

static_assert(sizeof(size_t) == 8, "the archive stores sizes as 8-byte size_t");

namespace {

// component types (component_type.cpp:23-38)
enum CT { kB = 0, kS32, kS64, kF32, kF64, kCTCount };
const char *const kCTName[kCTCount] = {"b", "s32", "s64", "f32", "f64"};
const size_t kCTBytes[kCTCount] = {1, 4, 8, 4, 8};

using Shape = std::vector<uint64_t>;

uint64_t extent_of(const Shape &s) {
  uint64_t n = 1;
  for (uint64_t e : s)
    n *= e;
  return n;
}

Shape shape_of_components(int comps) {
  if (comps == 1)
    return {};
  if (comps == 3)
    return {3};
  if (comps == 9)
    return {3, 3};
  throw std::logic_error("field with an unexpected component count");
}

// ---------------------------------------------------------------- writing
class FileOut {
 public:
  explicit FileOut(const char *path) : f_(std::fopen(path, "wb")) {
    if (!f_)
      throw std::invalid_argument(std::string("archive: cannot open for writing: ") + path);
  }
  ~FileOut() {
    if (f_)
      std::fclose(f_);
  }
  void bytes(const void *p, size_t n) {
    if (n && std::fwrite(p, 1, n, f_) != n)
      throw std::runtime_error("archive: write failed");
    written_ += n;
  }
  void size(uint64_t v) { bytes(&v, 8); }
  void str(std::string_view s) {
    size(s.size());
    bytes(s.data(), s.size());
  }
  void type(int ct, const Shape &shape) {
    str(kCTName[ct]);
    size(shape.size());
    for (uint64_t e : shape)
      size(e);
  }
  void close() {
    FILE *f = f_;
    f_ = nullptr;
    if (std::fclose(f) != 0)
      throw std::runtime_error("archive: close failed");
  }
  uint64_t written() const { return written_; }

 private:
  FILE *f_;
  uint64_t written_ = 0;
};

// one uniform field: type and the item's component bytes
struct UniformOut {
  int ct;
  Shape shape;
  std::vector<unsigned char> bytes;
};

template <class V>
UniformOut number(int ct, std::initializer_list<V> vals) {
  UniformOut u{ct, vals.size() == 1 ? Shape{} : Shape{vals.size()}, {}};
  for (V v : vals) {
    unsigned char b[8];
    switch (ct) {
      case kS32: { const int32_t x = static_cast<int32_t>(v); std::memcpy(b, &x, 4); break; }
      case kS64: { const int64_t x = static_cast<int64_t>(v); std::memcpy(b, &x, 8); break; }
      case kF32: { const float x = static_cast<float>(v); std::memcpy(b, &x, 4); break; }
      case kF64: { const double x = static_cast<double>(v); std::memcpy(b, &x, 8); break; }
      default: throw std::logic_error("unsupported uniform type");
    }
    u.bytes.insert(u.bytes.end(), b, b + kCTBytes[ct]);
  }
  return u;
}

// one varying field: a device field, or host values (the slots)
struct VaryingOut {
  int ct;
  Shape shape;
  int field = -1;
  std::vector<int32_t> host;
};

struct GroupOut {
  std::string type;
  std::vector<std::string> tags;
  int gid;
  uint64_t items;
  std::map<std::string, VaryingOut> varying;
  std::map<std::string, UniformOut> uniform;
};

void write_uniforms(FileOut &o, const std::map<std::string, UniformOut> &m) {
  o.size(m.size());
  for (const auto &[name, u] : m) {
    o.str(name);
    o.type(u.ct, u.shape);
    o.bytes(u.bytes.data(), u.bytes.size());
  }
}

struct PinnedPair {
  void *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  explicit PinnedPair(size_t bytes) {
    for (int k = 0; k < 2; ++k) {
      PRTCL_HIP_CHECK(hipHostMalloc(&buf[k], std::max<size_t>(bytes, 64), hipHostMallocDefault));
      PRTCL_HIP_CHECK(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
    }
  }
  ~PinnedPair() {
    for (int k = 0; k < 2; ++k) {
      if (ev[k])
        (void)hipEventDestroy(ev[k]);
      if (buf[k])
        (void)hipHostFree(buf[k]);
    }
  }
};

}  // namespace

uint64_t save_scene_archive(SceneBase &sc, CtxResources &res, const char *path) {
  const prtcl_hip_scene_desc &d = sc.desc;
  const int rt = d.real_type == 1 ? kF64 : kF32;
  const size_t rb = kCTBytes[rt];

  // global fields of the schemes on the path (iisph.hpp:136-142,
  // symplectic_euler, gravity), then the device state
  std::map<std::string, UniformOut> globals;
  globals["smoothing_scale"] = number(rt, {d.smoothing_scale});
  globals["time_step"] = number(rt, {d.time_step});
  globals["fade_duration"] = number(rt, {d.fade_duration});
  globals["iisph_relaxation"] = number(rt, {d.iisph_relaxation});
  globals["gravity"] = number(rt, {d.gravity[0], d.gravity[1], d.gravity[2]});
  globals["current_time"] = number(rt, {sc.get_global(PRTCL_HIP_G_CURRENT_TIME)});
  globals["maximum_speed"] = number(rt, {sc.get_global(PRTCL_HIP_G_MAXIMUM_SPEED)});
  globals["iisph_aprde"] = number(rt, {sc.get_global(PRTCL_HIP_G_IISPH_APRDE)});
  globals["iisph_nprde"] = number(kS32, {sc.get_global(PRTCL_HIP_G_IISPH_NPRDE)});
  globals["device_clock"] = number(kF64, {sc.exact_clock()});
  globals["device_steps_since_permute"] = number(kS32, {sc.steps_since_permute()});

  std::map<std::string, GroupOut> groups;
  const struct {
    const char *name, *type;
    int gid;
    uint64_t items;
  } gdefs[2] = {{"b", "boundary", PRTCL_HIP_BOUNDARY, static_cast<uint64_t>(d.n_boundary)},
                {"f", "fluid", PRTCL_HIP_FLUID, static_cast<uint64_t>(d.n_fluid)}};
  for (const auto &gd : gdefs) {
    GroupOut g{gd.type, {}, gd.gid, gd.items, {}, {}};
    if (gd.gid == PRTCL_HIP_FLUID) {
      g.tags = {"dynamic"};
      g.uniform["rest_density"] = number(rt, {d.rest_density});
      g.uniform["dynamic_viscosity"] = number(rt, {d.fluid_viscosity});
      if (d.implicit_viscosity == 1 || d.implicit_viscosity == 2) {
        g.uniform["wkbb18_maximum_error"] = number(rt, {d.wkbb18_maximum_error});
        g.uniform["wkbb18_maximum_iterations"] = number(kS32, {d.wkbb18_maximum_iterations});
        g.uniform["wkbb18_iterations"] =
            number(kS32, {sc.get_global(PRTCL_HIP_G_WKBB18_ITERATIONS)});
      }
      if (d.implicit_viscosity == 3) {
        g.uniform["strain_rate_viscosity"] =
            number(rt, {sc.get_global(PRTCL_HIP_G_STRAIN_RATE_VISCOSITY)});
        g.uniform["pt16_maximum_error"] = number(rt, {sc.get_global(PRTCL_HIP_G_PT16_MAXIMUM_ERROR)});
        g.uniform["pt16_maximum_iterations"] =
            number(kS32, {sc.get_global(PRTCL_HIP_G_PT16_MAXIMUM_ITERATIONS)});
        g.uniform["pt16_iterations"] = number(kS32, {sc.get_global(PRTCL_HIP_G_PT16_ITERATIONS)});
      }
    } else {
      g.uniform["dynamic_viscosity"] = number(rt, {d.boundary_viscosity});
    }
    for (int f = 0; f < PRTCL_HIP_F_COUNT; ++f) {
      const int comps = sc.field_components(gd.gid, f);
      if (comps > 0)
        g.varying[field_name(f)] = VaryingOut{rt, shape_of_components(comps), f, {}};
    }
    VaryingOut slot{kS32, {}, -1, std::vector<int32_t>(gd.items)};
    if (gd.items > 0)
      sc.slots(gd.gid, slot.host.data());
    g.varying["device_slot"] = std::move(slot);
    groups[gd.name] = std::move(g);
  }

  // device fields in file order, and the largest of them
  struct Pending {
    int gid, field;
    size_t count;
  };
  std::vector<Pending> dev;
  size_t maxb = 0;
  for (const auto &[gname, g] : groups)
    for (const auto &[fname, v] : g.varying)
      if (v.field >= 0 && g.items > 0) {
        const size_t count = g.items * extent_of(v.shape);
        dev.push_back({g.gid, v.field, count});
        maxb = std::max(maxb, count * rb);
      }
  PRTCL_HIP_CHECK(hipStreamSynchronize(res.stream));
  if (maxb > 0)
    res.ensure_staging(maxb);
  PinnedPair pin(maxb);
  size_t next = 0;
  auto enqueue = [&]() {
    if (next >= dev.size())
      return;
    const int b = static_cast<int>(next & 1);
    sc.download_async(dev[next].gid, dev[next].field, pin.buf[b], dev[next].count);
    PRTCL_HIP_CHECK(hipEventRecord(pin.ev[b], res.stream));
    ++next;
  };
  enqueue();
  enqueue();

  FileOut o(path);
  write_uniforms(o, globals);
  o.size(groups.size());
  size_t done = 0;
  for (const auto &[gname, g] : groups) {
    o.str(gname);
    o.str(g.type);
    o.size(g.tags.size());
    for (const std::string &t : g.tags)
      o.str(t);
    o.size(g.items);
    o.size(g.varying.size());
    for (const auto &[fname, v] : g.varying) {
      o.str(fname);
      o.type(v.ct, v.shape);
      o.size(g.items);
      if (g.items == 0)
        continue;
      if (v.field < 0) {
        o.bytes(v.host.data(), v.host.size() * sizeof(int32_t));
        continue;
      }
      const int b = static_cast<int>(done & 1);
      PRTCL_HIP_CHECK(hipEventSynchronize(pin.ev[b]));
      o.bytes(pin.buf[b], dev[done].count * rb);
      ++done;
      enqueue();  // refills the buffer just written
    }
    write_uniforms(o, g.uniform);
  }
  o.close();
  return o.written();
}

// ---------------------------------------------------------------- reading
struct FieldIn {
  int ct = -1;
  Shape shape;
  uint64_t count = 0;   // items (varying fields)
  uint64_t offset = 0;  // file offset of the values (varying fields)
  std::vector<unsigned char> values;  // uniform fields

  bool number(size_t k, double *out) const {
    if (k >= extent_of(shape) || values.size() < (k + 1) * kCTBytes[ct])
      return false;
    const unsigned char *p = values.data() + k * kCTBytes[ct];
    switch (ct) {
      case kB: *out = p[0] ? 1.0 : 0.0; return true;
      case kS32: { int32_t x; std::memcpy(&x, p, 4); *out = x; return true; }
      case kS64: { int64_t x; std::memcpy(&x, p, 8); *out = static_cast<double>(x); return true; }
      case kF32: { float x; std::memcpy(&x, p, 4); *out = x; return true; }
      case kF64: { double x; std::memcpy(&x, p, 8); *out = x; return true; }
      default: return false;
    }
  }
};

struct GroupIn {
  std::string name, type;
  std::vector<std::string> tags;
  uint64_t items = 0;
  std::map<std::string, FieldIn> varying, uniform;
};

struct ParsedArchive {
  std::string path;
  std::map<std::string, FieldIn> globals;
  std::vector<GroupIn> groups;
  int fluid = -1, boundary = -1;  // indexes into groups
  int real_ct = kF32;
};

void ParsedArchiveDeleter::operator()(ParsedArchive *p) const { delete p; }

namespace {

class FileIn {
 public:
  explicit FileIn(const char *path) : f_(std::fopen(path, "rb")) {
    if (!f_)
      throw std::invalid_argument(std::string("archive: cannot open: ") + path);
    struct stat st;
    if (fstat(fileno(f_), &st) != 0)
      throw std::invalid_argument("archive: cannot stat the file");
    bytes_ = static_cast<uint64_t>(st.st_size);
  }
  ~FileIn() {
    if (f_)
      std::fclose(f_);
  }
  void read(void *p, size_t n) {
    if (n > bytes_ - pos_ || (n && std::fread(p, 1, n, f_) != n))
      throw std::invalid_argument("archive: unexpected end of data");
    pos_ += n;
  }
  uint64_t size() {
    uint64_t v;
    read(&v, 8);
    return v;
  }
  std::string str() {
    const uint64_t n = size();
    if (n > (uint64_t(1) << 20))
      throw std::invalid_argument("archive: string longer than 1 MiB");
    std::string s(n, '\0');
    read(s.data(), n);
    return s;
  }
  void type(FieldIn &f) {
    const std::string c = str();
    f.ct = -1;
    for (int k = 0; k < kCTCount; ++k)
      if (c == kCTName[k])
        f.ct = k;
    if (f.ct < 0)
      throw std::invalid_argument("archive: invalid component type \"" + c + "\"");
    const uint64_t rank = size();
    if (rank > 8)
      throw std::invalid_argument("archive: field of rank above 8");
    f.shape.resize(rank);
    for (uint64_t &e : f.shape)
      e = size();
    if (extent_of(f.shape) > (uint64_t(1) << 20))
      throw std::invalid_argument("archive: item of more than 2^20 components");
  }
  void skip(uint64_t n) {
    if (n > bytes_ - pos_)
      throw std::invalid_argument("archive: unexpected end of data");
    if (fseeko(f_, static_cast<off_t>(pos_ + n), SEEK_SET) != 0)
      throw std::invalid_argument("archive: seek failed");
    pos_ += n;
  }
  void seek(uint64_t at) {
    if (at > bytes_ || fseeko(f_, static_cast<off_t>(at), SEEK_SET) != 0)
      throw std::invalid_argument("archive: seek failed");
    pos_ = at;
  }
  uint64_t pos() const { return pos_; }
  uint64_t bytes() const { return bytes_; }

 private:
  FILE *f_;
  uint64_t bytes_ = 0, pos_ = 0;
};

void read_uniforms(FileIn &in, std::map<std::string, FieldIn> &out) {
  const uint64_t n = in.size();
  for (uint64_t k = 0; k < n; ++k) {
    const std::string name = in.str();
    FieldIn f;
    in.type(f);
    f.values.resize(extent_of(f.shape) * kCTBytes[f.ct]);
    in.read(f.values.data(), f.values.size());
    out[name] = std::move(f);
  }
}

bool num(const std::map<std::string, FieldIn> &m, const char *name, double *out, size_t k = 0) {
  auto it = m.find(name);
  return it != m.end() && it->second.ct != kB && it->second.number(k, out);
}

template <class V>
void num_into(const std::map<std::string, FieldIn> &m, const char *name, V *out) {
  double v;
  if (num(m, name, &v))
    *out = static_cast<V>(v);
}

}  // namespace

ParsedArchivePtr parse_scene_archive(const char *path, prtcl_hip_scene_desc &desc) {
  ParsedArchivePtr a(new ParsedArchive);
  a->path = path;
  FileIn in(path);
  read_uniforms(in, a->globals);
  const uint64_t ng = in.size();
  if (ng > 4096)
    throw std::invalid_argument("archive: more than 4096 groups");
  for (uint64_t k = 0; k < ng; ++k) {
    GroupIn g;
    g.name = in.str();
    g.type = in.str();
    const uint64_t nt = in.size();
    if (nt > 4096)
      throw std::invalid_argument("archive: more than 4096 tags");
    for (uint64_t t = 0; t < nt; ++t)
      g.tags.push_back(in.str());
    g.items = in.size();
    const uint64_t nf = in.size();
    for (uint64_t t = 0; t < nf; ++t) {
      const std::string name = in.str();
      FieldIn f;
      in.type(f);
      f.count = in.size();
      f.offset = in.pos();
      const uint64_t comps = extent_of(f.shape);
      if (f.count > (uint64_t(1) << 31) || f.count * comps > (in.bytes() - in.pos()) / kCTBytes[f.ct])
        throw std::invalid_argument("archive: unexpected end of data");
      in.skip(f.count * comps * kCTBytes[f.ct]);
      g.varying[name] = std::move(f);
    }
    read_uniforms(in, g.uniform);
    const int idx = static_cast<int>(a->groups.size());
    if (g.type == "fluid" || g.type == "boundary") {
      int &slot = g.type == "fluid" ? a->fluid : a->boundary;
      if (slot >= 0)
        throw std::invalid_argument("archive: a scene holds one group of type " + g.type);
      slot = idx;
    }
    a->groups.push_back(std::move(g));
  }
  if (in.pos() != in.bytes())
    throw std::invalid_argument("archive: data after the model");

  // counts, real type, parameters
  const GroupIn *fl = a->fluid >= 0 ? &a->groups[a->fluid] : nullptr;
  const GroupIn *bd = a->boundary >= 0 ? &a->groups[a->boundary] : nullptr;
  desc.n_fluid = fl ? static_cast<int64_t>(fl->items) : 0;
  desc.n_boundary = bd ? static_cast<int64_t>(bd->items) : 0;
  for (const GroupIn *g : {bd, fl}) {
    if (!g)
      continue;
    auto it = g->varying.find("position");
    if (it != g->varying.end() && (it->second.ct == kF32 || it->second.ct == kF64))
      a->real_ct = it->second.ct;
  }
  desc.real_type = a->real_ct == kF64 ? 1 : 0;
  num_into(a->globals, "smoothing_scale", &desc.smoothing_scale);
  num_into(a->globals, "time_step", &desc.time_step);
  num_into(a->globals, "fade_duration", &desc.fade_duration);
  num_into(a->globals, "iisph_relaxation", &desc.iisph_relaxation);
  for (int k = 0; k < 3; ++k) {
    double v;
    if (num(a->globals, "gravity", &v, k))
      desc.gravity[k] = v;
  }
  if (fl) {
    num_into(fl->uniform, "rest_density", &desc.rest_density);
    num_into(fl->uniform, "dynamic_viscosity", &desc.fluid_viscosity);
    num_into(fl->uniform, "wkbb18_maximum_error", &desc.wkbb18_maximum_error);
    num_into(fl->uniform, "wkbb18_maximum_iterations", &desc.wkbb18_maximum_iterations);
    num_into(fl->uniform, "strain_rate_viscosity", &desc.pt16_strain_rate_viscosity);
    if (fl->varying.count("wkbb18_system_diagonal"))
      desc.implicit_viscosity = fl->varying.count("gradient_correction") ? 2 : 1;
    if (fl->varying.count("pt16_same_density"))
      desc.implicit_viscosity = 3;
  }
  if (bd)
    num_into(bd->uniform, "dynamic_viscosity", &desc.boundary_viscosity);
  return a;
}

uint64_t load_scene_archive(SceneBase &sc, const ParsedArchive &a) {
  FileIn in(a.path.c_str());
  const size_t rb = kCTBytes[a.real_ct];
  uint64_t loaded = 0;
  // the largest varying field the scene takes
  size_t maxb = 4;
  for (const GroupIn &g : a.groups)
    for (const auto &[name, f] : g.varying)
      maxb = std::max<size_t>(maxb, f.count * extent_of(f.shape) * kCTBytes[f.ct]);
  void *buf = nullptr;
  PRTCL_HIP_CHECK(hipHostMalloc(&buf, maxb, hipHostMallocDefault));
  struct Free {
    void *p;
    ~Free() { (void)hipHostFree(p); }
  } free_buf{buf};

  const struct {
    int index, gid;
  } gs[2] = {{a.boundary, PRTCL_HIP_BOUNDARY}, {a.fluid, PRTCL_HIP_FLUID}};
  for (const auto &gd : gs) {
    if (gd.index < 0)
      continue;
    const GroupIn &g = a.groups[gd.index];
    if (g.items == 0)
      continue;
    // slots before any field, so every field lands in its saved slot
    auto it = g.varying.find("device_slot");
    if (it != g.varying.end()) {
      const FieldIn &f = it->second;
      if (f.ct != kS32 || !f.shape.empty() || f.count != g.items)
        throw std::invalid_argument("archive: device_slot of group " + g.name +
                                    " is not one s32 per particle");
      in.seek(f.offset);
      in.read(buf, 4 * f.count);
      sc.set_slots(gd.gid, static_cast<const int32_t *>(buf));
      loaded += 4 * f.count;
    }
    for (const auto &[name, f] : g.varying) {
      const int id = prtcl_hip_field_id(name.c_str());
      if (id < 0)
        continue;  // a field of another scheme, or the slots
      const int comps = sc.field_components(gd.gid, id);
      if (comps == 0)
        continue;  // a field this scene's schemes do not bind
      if (f.ct != a.real_ct || extent_of(f.shape) != static_cast<uint64_t>(comps))
        throw std::invalid_argument("archive: field " + name + " of group " + g.name +
                                    " has a different type");
      if (f.count != g.items)
        throw std::invalid_argument("archive: field " + name + " of group " + g.name +
                                    " has a different item count");
      const size_t count = f.count * comps;
      in.seek(f.offset);
      in.read(buf, count * rb);
      sc.upload(gd.gid, id, buf, count);
      loaded += count * rb;
    }
  }

  double v;
  if (num(a.globals, "device_clock", &v))
    sc.set_exact_clock(v);
  else if (num(a.globals, "current_time", &v))
    sc.set_exact_clock(v);
  if (num(a.globals, "device_steps_since_permute", &v))
    sc.set_steps_since_permute(static_cast<int>(v));
  if (num(a.globals, "maximum_speed", &v))
    sc.set_global(PRTCL_HIP_G_MAXIMUM_SPEED, v);
  if (a.fluid >= 0 && sc.desc.implicit_viscosity == 3) {
    const auto &u = a.groups[a.fluid].uniform;
    if (num(u, "pt16_maximum_error", &v))
      sc.set_global(PRTCL_HIP_G_PT16_MAXIMUM_ERROR, v);
    if (num(u, "pt16_maximum_iterations", &v))
      sc.set_global(PRTCL_HIP_G_PT16_MAXIMUM_ITERATIONS, v);
    if (num(u, "pt16_iterations", &v))
      sc.set_global(PRTCL_HIP_G_PT16_ITERATIONS, v);
  }
  return loaded;
}

}  // namespace prtcl_amd
