// dat.hpp -- the reference's serial .dat layout (host code, no HIP).
#pragma once

#include <cstddef>
#include <functional>
#include <string>

#include "../../include/mmsp_amd.h"

namespace mmspx {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// `node_major` holds nodes in row-major order with each node's fields
// contiguous; elem is sizeof(double) or sizeof(float).
int dat_write(const char* filename, const mmspx_desc& d, const void* node_major, size_t elem, int compress);
int dat_read_header(const char* filename, mmspx_desc* d, int* file_f64);
int dat_read(const char* filename, const mmspx_desc& d, int file_f64, void* node_major);
// The same layout with the zlib stream deflated by `threads` host threads in
// chunks (one zlib stream, one block). Before reading node-major bytes [0, n)
// a worker calls wait_bytes(n), which returns false if they will never come.
int dat_write_par(const char* filename, const mmspx_desc& d, const void* node_major, size_t elem, int compress,
                  int threads, const std::function<bool(size_t)>& wait_bytes);

}  // namespace mmspx
