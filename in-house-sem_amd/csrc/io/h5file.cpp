#include "h5file.hpp"

#include <hdf5.h>

#include <type_traits>

#include "../host/error.hpp"

namespace insem_amd::h5 {
namespace {

// An HDF5 identifier closed with its own close function when it leaves scope.
struct Hid {
  hid_t v = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid(hid_t id, herr_t (*c)(hid_t)) : v(id), close(c) {}
  Hid(const Hid&) = delete;
  Hid& operator=(const Hid&) = delete;
  ~Hid() {
    if (v >= 0 && close) close(v);
  }
  bool ok() const { return v >= 0; }
};

template <class T>
hid_t mem_type() {
  if constexpr (std::is_same_v<T, double>) return H5T_NATIVE_DOUBLE;
  if constexpr (std::is_same_v<T, float>) return H5T_NATIVE_FLOAT;
  if constexpr (std::is_same_v<T, int32_t>) return H5T_NATIVE_INT32;
  if constexpr (std::is_same_v<T, int64_t>) return H5T_NATIVE_INT64;
}

// The components of an absolute path: "/a/b/c" -> {"/a", "/a/b", "/a/b/c"}.
std::vector<std::string> prefixes(const std::string& path) {
  INSEM_REQUIRE(!path.empty() && path[0] == '/', "HDF5 object paths start with '/': '" << path << "'");
  std::vector<std::string> out;
  for (std::size_t i = 1; i <= path.size(); ++i)
    if (i == path.size() || path[i] == '/') {
      if (i > 1 && path[i - 1] != '/') out.push_back(path.substr(0, i));
    }
  return out;
}

std::string parent_of(const std::string& path) {
  const std::size_t k = path.rfind('/');
  return k == 0 || k == std::string::npos ? "/" : path.substr(0, k);
}

void quiet() {
  static const bool once = [] {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors surface as Error, not stderr dumps
    return true;
  }();
  (void)once;
}

}  // namespace

File::File(const std::string& path, Open how) : path_(path) {
  quiet();
  hid_t id = -1;
  if (how == Open::Create)
    id = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  else
    id = H5Fopen(path.c_str(), how == Open::Read ? H5F_ACC_RDONLY : H5F_ACC_RDWR, H5P_DEFAULT);
  INSEM_REQUIRE(id >= 0, "cannot " << (how == Open::Create ? "create" : "open") << " HDF5 file '" << path << "'");
  id_ = id;
}

File::~File() {
  if (id_ >= 0) H5Fclose(static_cast<hid_t>(id_));
}

bool File::has(const std::string& path) const {
  if (path == "/") return true;
  for (const std::string& p : prefixes(path))
    if (H5Lexists(static_cast<hid_t>(id_), p.c_str(), H5P_DEFAULT) <= 0) return false;
  return true;
}

void File::make_group(const std::string& path) {
  if (path == "/") return;
  for (const std::string& p : prefixes(path)) {
    if (H5Lexists(static_cast<hid_t>(id_), p.c_str(), H5P_DEFAULT) > 0) continue;
    Hid g(H5Gcreate2(static_cast<hid_t>(id_), p.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    INSEM_REQUIRE(g.ok(), "cannot create group '" << p << "' in '" << path_ << "'");
  }
}

template <class T>
void File::put(const std::string& path, const T* data, const std::vector<std::size_t>& shape) {
  const hid_t f = static_cast<hid_t>(id_);
  make_group(parent_of(path));
  if (has(path)) INSEM_REQUIRE(H5Ldelete(f, path.c_str(), H5P_DEFAULT) >= 0, "cannot replace '" << path << "'");
  std::vector<hsize_t> ext(shape.begin(), shape.end());
  std::size_t n = 1;
  for (std::size_t e : shape) n *= e;
  Hid space(H5Screate_simple(static_cast<int>(ext.size()), ext.data(), nullptr), H5Sclose);
  INSEM_REQUIRE(space.ok(), "cannot describe the shape of '" << path << "'");
  Hid set(H5Dcreate2(f, path.c_str(), mem_type<T>(), space.v, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  INSEM_REQUIRE(set.ok(), "cannot create dataset '" << path << "' in '" << path_ << "'");
  if (n > 0)
    INSEM_REQUIRE(H5Dwrite(set.v, mem_type<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0,
                  "cannot write dataset '" << path << "' in '" << path_ << "'");
}

template <class T>
std::vector<T> File::get(const std::string& path, std::vector<std::size_t>* shape) const {
  INSEM_REQUIRE(has(path), "'" << path_ << "' has no dataset '" << path << "'");
  Hid set(H5Dopen2(static_cast<hid_t>(id_), path.c_str(), H5P_DEFAULT), H5Dclose);
  INSEM_REQUIRE(set.ok(), "'" << path << "' in '" << path_ << "' is not a dataset");
  Hid space(H5Dget_space(set.v), H5Sclose);
  const int rank = H5Sget_simple_extent_ndims(space.v);
  INSEM_REQUIRE(rank >= 0, "cannot read the shape of '" << path << "'");
  std::vector<hsize_t> ext(static_cast<std::size_t>(rank));
  if (rank > 0) H5Sget_simple_extent_dims(space.v, ext.data(), nullptr);
  std::size_t n = 1;
  for (hsize_t e : ext) n *= static_cast<std::size_t>(e);
  if (shape) shape->assign(ext.begin(), ext.end());
  std::vector<T> out(n);
  if (n > 0)
    INSEM_REQUIRE(H5Dread(set.v, mem_type<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0,
                  "cannot read dataset '" << path << "' of '" << path_ << "'");
  return out;
}

namespace {

// Writes (replacing) a scalar attribute of the given file type.
void put_attr(hid_t file, const std::string& fname, const std::string& object, const std::string& name,
              hid_t type, const void* value) {
  Hid obj(H5Oopen(file, object.c_str(), H5P_DEFAULT), H5Oclose);
  INSEM_REQUIRE(obj.ok(), "'" << fname << "' has no object '" << object << "' for attribute '" << name << "'");
  if (H5Aexists(obj.v, name.c_str()) > 0) H5Adelete(obj.v, name.c_str());
  Hid space(H5Screate(H5S_SCALAR), H5Sclose);
  Hid a(H5Acreate2(obj.v, name.c_str(), type, space.v, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  INSEM_REQUIRE(a.ok() && H5Awrite(a.v, type, value) >= 0,
                "cannot write attribute '" << name << "' of '" << object << "' in '" << fname << "'");
}

}  // namespace

// Strings are stored fixed-length and null-terminated.
void File::set(const std::string& object, const std::string& name, const std::string& value) {
  make_group(object);
  Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(t.v, value.size() + 1);
  H5Tset_strpad(t.v, H5T_STR_NULLTERM);
  put_attr(static_cast<hid_t>(id_), path_, object, name, t.v, value.c_str());
}

void File::set(const std::string& object, const std::string& name, int64_t value) {
  make_group(object);
  put_attr(static_cast<hid_t>(id_), path_, object, name, H5T_NATIVE_INT64, &value);
}

void File::set(const std::string& object, const std::string& name, double value) {
  make_group(object);
  put_attr(static_cast<hid_t>(id_), path_, object, name, H5T_NATIVE_DOUBLE, &value);
}

bool File::has_attr(const std::string& object, const std::string& name) const {
  return has(object) && H5Aexists_by_name(static_cast<hid_t>(id_), object.c_str(), name.c_str(), H5P_DEFAULT) > 0;
}

namespace {

hid_t open_attr(hid_t file, const std::string& fname, const std::string& object, const std::string& name) {
  const hid_t a = H5Aopen_by_name(file, object.c_str(), name.c_str(), H5P_DEFAULT, H5P_DEFAULT);
  INSEM_REQUIRE(a >= 0, "'" << fname << "' has no attribute '" << name << "' on '" << object << "'");
  return a;
}

}  // namespace

char File::attr_class(const std::string& object, const std::string& name) const {
  Hid a(open_attr(static_cast<hid_t>(id_), path_, object, name), H5Aclose);
  Hid t(H5Aget_type(a.v), H5Tclose);
  switch (H5Tget_class(t.v)) {
    case H5T_STRING: return 's';
    case H5T_INTEGER: return 'i';
    case H5T_FLOAT: return 'f';
    default: return '?';
  }
}

std::string File::get_string(const std::string& object, const std::string& name) const {
  Hid a(open_attr(static_cast<hid_t>(id_), path_, object, name), H5Aclose);
  Hid ft(H5Aget_type(a.v), H5Tclose);
  INSEM_REQUIRE(H5Tget_class(ft.v) == H5T_STRING, "attribute '" << name << "' of '" << object << "' is not text");
  if (H5Tis_variable_str(ft.v) > 0) {
    Hid mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt.v, H5T_VARIABLE);
    char* p = nullptr;
    INSEM_REQUIRE(H5Aread(a.v, mt.v, &p) >= 0, "cannot read attribute '" << name << "' of '" << object << "'");
    std::string s = p ? p : "";
    H5free_memory(p);
    return s;
  }
  const std::size_t n = H5Tget_size(ft.v);
  std::string s(n, '\0');
  INSEM_REQUIRE(H5Aread(a.v, ft.v, &s[0]) >= 0, "cannot read attribute '" << name << "' of '" << object << "'");
  const std::size_t z = s.find('\0');
  if (z != std::string::npos) s.resize(z);
  return s;
}

int64_t File::get_int(const std::string& object, const std::string& name) const {
  Hid a(open_attr(static_cast<hid_t>(id_), path_, object, name), H5Aclose);
  int64_t v = 0;
  INSEM_REQUIRE(H5Aread(a.v, H5T_NATIVE_INT64, &v) >= 0, "cannot read attribute '" << name << "' of '" << object << "'");
  return v;
}

double File::get_double(const std::string& object, const std::string& name) const {
  Hid a(open_attr(static_cast<hid_t>(id_), path_, object, name), H5Aclose);
  double v = 0.0;
  INSEM_REQUIRE(H5Aread(a.v, H5T_NATIVE_DOUBLE, &v) >= 0, "cannot read attribute '" << name << "' of '" << object << "'");
  return v;
}

template void File::put<double>(const std::string&, const double*, const std::vector<std::size_t>&);
template void File::put<float>(const std::string&, const float*, const std::vector<std::size_t>&);
template void File::put<int32_t>(const std::string&, const int32_t*, const std::vector<std::size_t>&);
template void File::put<int64_t>(const std::string&, const int64_t*, const std::vector<std::size_t>&);
template std::vector<double> File::get<double>(const std::string&, std::vector<std::size_t>*) const;
template std::vector<float> File::get<float>(const std::string&, std::vector<std::size_t>*) const;
template std::vector<int32_t> File::get<int32_t>(const std::string&, std::vector<std::size_t>*) const;
template std::vector<int64_t> File::get<int64_t>(const std::string&, std::vector<std::size_t>*) const;

}  // namespace insem_amd::h5
