// HDF5 snapshot writer (the reference's save_instant_field format).
#pragma once
#include <cstdint>
#include <string>

namespace owt {

// One HDF5 file, created (truncated) on construction, closed on close() or
// destruction. Errors throw owt::Error with code 5 (I/O).
class H5Writer {
  public:
    explicit H5Writer(const std::string &path);
    ~H5Writer();
    H5Writer(const H5Writer &) = delete;
    H5Writer &operator=(const H5Writer &) = delete;
    // a [dims[0]][dims[1]][dims[2]] double dataset, row-major, last index fastest
    void write(const std::string &name, const long dims[3], const double *data);
    void close();

  private:
    std::string path_;
    int64_t file_ = -1;
};

}  // namespace owt
