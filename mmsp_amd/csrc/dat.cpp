// dat.cpp -- the reference's serial .dat checkpoint layout, written and read
// by the library so that state leaving the device stays readable by the
// reference's own readers (grid::input, mmsp2png, mmsp2vti, mmsp2tsv).
//
// Layout (include/MMSP.grid.cpp:1358-1402 output, 1818-1909 write_buffer,
// 1183-1356 input/read; names MMSP.grid.h:593-596, MMSP.vector.h:115-117):
//   "grid:<type>\n" dim "\n" fields "\n" ("g0 g1\n") x dim (dx "\n") x dim
//   int32 nblocks
//   per block: int32 (x0,x1) x dim, int32 (b0,b1) x dim, uint64 size_in_mem,
//              uint64 size_on_disk, payload
// The payload is nodes in row-major order; a vector<T> node is int32 n then n
// values (MMSP.vector.cpp:100-105). It is zlib compress2 level 9 unless the
// two sizes are equal, in which case it is raw (accepted by the reference's
// reader, MMSP.grid.cpp:1306-1330).
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "dat.hpp"

namespace mmspx {

static std::string type_name(const mmspx_desc& d, size_t elem)
{
	std::string t = elem == 8 ? "double" : "float";
	return d.vector ? "grid:vector:" + t : "grid:" + t;
}

int dat_write(const char* filename, const mmspx_desc& d, const void* host, size_t elem, int compress)
{
	const int q = d.vector ? d.fields : 1;
	int64_t nodes = 1;
	for (int k = 0; k < d.dim; k++) nodes *= d.n[k];

	std::ostringstream hdr;  // default precision 6, as the reference's stringstream
	hdr << type_name(d, elem) << '\n' << d.dim << '\n' << d.fields << '\n';
	for (int k = 0; k < d.dim; k++) hdr << d.g0[k] << " " << d.g0[k] + d.n[k] << '\n';
	for (int k = 0; k < d.dim; k++) hdr << d.dx[k] << '\n';

	// raw payload
	const size_t node_bytes = d.vector ? sizeof(int32_t) + q * elem : elem;
	std::vector<char> raw((size_t)nodes * node_bytes);
	if (!d.vector) {
		std::memcpy(raw.data(), host, raw.size());
	} else {
		const char* src = (const char*)host;
		char* dst = raw.data();
		const int32_t qq = q;
		for (int64_t i = 0; i < nodes; i++) {
			std::memcpy(dst, &qq, sizeof(qq));
			std::memcpy(dst + sizeof(qq), src + i * q * elem, q * elem);
			dst += node_bytes;
		}
	}
	uint64_t size_in_mem = raw.size();
	std::vector<char> packed;
	const char* body = raw.data();
	uint64_t size_on_disk = size_in_mem;
	if (compress) {
		uLongf cap = compressBound((uLong)size_in_mem);
		packed.resize(cap);
		int zs = compress2((Bytef*)packed.data(), &cap, (const Bytef*)raw.data(), (uLong)size_in_mem, 9);
		if (zs != Z_OK) return fail(MMSPX_EIO, "zlib compress2 failed");
		size_on_disk = cap;
		body = packed.data();
		// A compressed block exactly the raw size would read back as raw.
		if (size_on_disk == size_in_mem) {
			body = raw.data();
		}
	}

	std::ofstream f(filename, std::ios::binary);
	if (!f) return fail(MMSPX_EIO, std::string("could not open ") + filename);
	const std::string h = hdr.str();
	f.write(h.data(), (std::streamsize)h.size());
	int32_t nb = 1;
	f.write((const char*)&nb, sizeof(nb));
	for (int k = 0; k < d.dim; k++) {
		int32_t lo = d.g0[k], hi = d.g0[k] + d.n[k];
		f.write((const char*)&lo, 4);
		f.write((const char*)&hi, 4);
	}
	for (int k = 0; k < d.dim; k++) {
		int32_t lo = d.b0[k], hi = d.b1[k];
		f.write((const char*)&lo, 4);
		f.write((const char*)&hi, 4);
	}
	f.write((const char*)&size_in_mem, 8);
	f.write((const char*)&size_on_disk, 8);
	f.write(body, (std::streamsize)size_on_disk);
	if (!f) return fail(MMSPX_EIO, std::string("write failed: ") + filename);
	return 0;
}

// Parses the text header; leaves `in` at the block count.
static int parse_header(std::ifstream& in, const char* filename, mmspx_desc* d, int* file_f64)
{
	std::string type;
	std::getline(in, type, '\n');
	if (type.compare(0, 5, "grid:") != 0)
		return fail(MMSPX_EFORMAT, std::string(filename) + ": file does not contain grid data");
	std::memset(d, 0, sizeof(*d));
	if (type == "grid:double" || type == "grid:scalar:double") {
		*file_f64 = 1;
	} else if (type == "grid:float" || type == "grid:scalar:float") {
		*file_f64 = 0;
	} else if (type == "grid:vector:double") {
		*file_f64 = 1;
		d->vector = 1;
	} else if (type == "grid:vector:float") {
		*file_f64 = 0;
		d->vector = 1;
	} else {
		return fail(MMSPX_EFORMAT, std::string(filename) + ": unsupported grid type " + type);
	}
	in >> d->dim >> d->fields;
	if (!in || d->dim < 1 || d->dim > 3) return fail(MMSPX_EFORMAT, std::string(filename) + ": bad dimension");
	for (int k = 0; k < 3; k++) {
		d->n[k] = 1;
		d->dx[k] = 1.0;
		d->b0[k] = d->b1[k] = MMSPX_PERIODIC;
	}
	for (int k = 0; k < d->dim; k++) {
		int lo, hi;
		in >> lo >> hi;
		d->g0[k] = lo;
		d->n[k] = hi - lo;
	}
	for (int k = 0; k < d->dim; k++) in >> d->dx[k];
	in.ignore(10, '\n');
	if (!in) return fail(MMSPX_EFORMAT, std::string(filename) + ": truncated header");
	d->precision = *file_f64 ? MMSPX_F64 : MMSPX_F32;
	return 0;
}

int dat_read_header(const char* filename, mmspx_desc* d, int* file_f64)
{
	std::ifstream in(filename, std::ios::binary | std::ios::ate);
	if (!in) return fail(MMSPX_EIO, std::string("could not open ") + filename);
	const std::streamoff file_size = in.tellg();
	in.seekg(0);
	int rc = parse_header(in, filename, d, file_f64);
	if (rc) return rc;
	// boundary conditions come from the blocks that touch the global limits
	int32_t nb = 0;
	in.read((char*)&nb, 4);
	for (int b = 0; b < nb && in; b++) {
		int32_t lim[6], bc[6];
		in.read((char*)lim, 8 * d->dim);
		in.read((char*)bc, 8 * d->dim);
		uint64_t sm = 0, sd = 0;
		in.read((char*)&sm, 8);
		in.read((char*)&sd, 8);
		for (int k = 0; k < d->dim; k++) {
			if (lim[2 * k] == d->g0[k]) d->b0[k] = bc[2 * k];
			if (lim[2 * k + 1] == d->g0[k] + d->n[k]) d->b1[k] = bc[2 * k + 1];
		}
		if (!in || (std::streamoff)in.tellg() + (std::streamoff)sd > file_size)
			return fail(MMSPX_EFORMAT, std::string(filename) + ": truncated block data");
		in.seekg((std::streamoff)sd, std::ios::cur);
	}
	if (!in) return fail(MMSPX_EFORMAT, std::string(filename) + ": truncated block table");
	return 0;
}

int dat_read(const char* filename, const mmspx_desc& want, int file_f64, void* host)
{
	std::ifstream in(filename, std::ios::binary);
	if (!in) return fail(MMSPX_EIO, std::string("could not open ") + filename);
	mmspx_desc d;
	int f64 = 0;
	int rc = parse_header(in, filename, &d, &f64);
	if (rc) return rc;
	// `host` was sized from `want`: every extent, origin and the field count
	// must still match, or a block could be scattered past its end
	bool same = d.dim == want.dim && f64 == file_f64 && d.vector == want.vector &&
	            (d.vector ? d.fields : 1) == (want.vector ? want.fields : 1);
	for (int k = 0; k < d.dim && same; k++) same = d.n[k] == want.n[k] && d.g0[k] == want.g0[k];
	if (!same) return fail(MMSPX_EFORMAT, std::string(filename) + ": header changed while reading");
	const size_t elem = f64 ? 8 : 4;
	const int q = d.vector ? d.fields : 1;
	const size_t node_bytes = d.vector ? 4 + q * elem : elem;
	int32_t nb = 0;
	in.read((char*)&nb, 4);
	char* out = (char*)host;
	for (int b = 0; b < nb; b++) {
		int32_t lim[6], bc[6];
		in.read((char*)lim, 8 * d.dim);
		in.read((char*)bc, 8 * d.dim);
		uint64_t size_in_mem = 0, size_on_disk = 0;
		in.read((char*)&size_in_mem, 8);
		in.read((char*)&size_on_disk, 8);
		if (!in) return fail(MMSPX_EFORMAT, std::string(filename) + ": truncated block header");
		int ext[3] = {1, 1, 1}, lo[3] = {0, 0, 0};
		int64_t bn = 1;
		for (int k = 0; k < d.dim; k++) {
			lo[k] = lim[2 * k] - d.g0[k];
			ext[k] = lim[2 * k + 1] - lim[2 * k];
			if (lo[k] < 0 || ext[k] < 0 || lo[k] + ext[k] > d.n[k])
				return fail(MMSPX_EFORMAT, std::string(filename) + ": block outside the grid");
			bn *= ext[k];
		}
		if (size_in_mem != (uint64_t)bn * node_bytes)
			return fail(MMSPX_EFORMAT, std::string(filename) + ": block size does not match its limits");
		std::vector<char> body(size_on_disk);
		in.read(body.data(), (std::streamsize)size_on_disk);
		if (!in) return fail(MMSPX_EFORMAT, std::string(filename) + ": truncated block data");
		std::vector<char> raw;
		const char* src = body.data();
		if (size_in_mem != size_on_disk) {
			raw.resize(size_in_mem);
			uLongf got = (uLongf)size_in_mem;
			if (uncompress((Bytef*)raw.data(), &got, (const Bytef*)body.data(), (uLong)size_on_disk) != Z_OK ||
			    got != size_in_mem)
				return fail(MMSPX_EFORMAT, std::string(filename) + ": zlib uncompress failed");
			src = raw.data();
		}
		// scatter the block into the global node-major array
		int64_t strides[3] = {1, 1, 1};
		for (int k = d.dim - 2; k >= 0; k--) strides[k] = strides[k + 1] * d.n[k + 1];
		for (int64_t i = 0; i < bn; i++) {
			int64_t r = i, gi = 0;
			for (int k = d.dim - 1; k >= 0; k--) {
				int64_t c = r % ext[k];
				r /= ext[k];
				gi += (c + lo[k]) * strides[k];
			}
			const char* s = src + i * node_bytes;
			if (d.vector) {
				int32_t nq;
				std::memcpy(&nq, s, 4);
				if (nq != q) return fail(MMSPX_EFORMAT, std::string(filename) + ": node field count mismatch");
				s += 4;
			}
			std::memcpy(out + gi * q * elem, s, q * elem);
		}
	}
	return 0;
}

}  // namespace mmspx
