// HDF5 access for case and results files, over the HDF5 C library.
//
// A file is a tree of groups; datasets hold n-dimensional arrays of
// double, float, int32 or int64 and objects carry scalar attributes (string,
// int64, double).  Paths are absolute ("/mesh/nodes").  Writing a dataset or
// attribute creates missing parent groups and replaces an existing one of
// the same name, so a file can be rewritten piecewise.  Every failure throws
// insem_amd::Error naming the file and the object.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace insem_amd::h5 {

enum class Open { Read, Update, Create };

class File {
 public:
  File(const std::string& path, Open how);
  ~File();
  File(const File&) = delete;
  File& operator=(const File&) = delete;

  const std::string& path() const { return path_; }

  // Whether every link of the path exists ("/" always does).
  bool has(const std::string& path) const;
  void make_group(const std::string& path);

  template <class T>
  void put(const std::string& path, const T* data, const std::vector<std::size_t>& shape);
  template <class T>
  void put(const std::string& path, const std::vector<T>& v) {
    put(path, v.data(), {v.size()});
  }
  // The whole dataset, flattened in C order; *shape (optional) its extents.
  template <class T>
  std::vector<T> get(const std::string& path, std::vector<std::size_t>* shape = nullptr) const;

  void set(const std::string& object, const std::string& name, const std::string& value);
  void set(const std::string& object, const std::string& name, const char* value) {
    set(object, name, std::string(value));
  }
  void set(const std::string& object, const std::string& name, int64_t value);
  void set(const std::string& object, const std::string& name, double value);
  bool has_attr(const std::string& object, const std::string& name) const;
  // Readers convert what the file stores to the asked type (HDF5's
  // conversion rules); strings may be fixed-length or variable-length.
  // 's' text, 'i' integer, 'f' floating point, '?' anything else
  char attr_class(const std::string& object, const std::string& name) const;
  std::string get_string(const std::string& object, const std::string& name) const;
  int64_t get_int(const std::string& object, const std::string& name) const;
  double get_double(const std::string& object, const std::string& name) const;

 private:
  int64_t id_ = -1;  // hid_t
  std::string path_;
};

}  // namespace insem_amd::h5
