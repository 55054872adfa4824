// dat_par.cpp -- the serial .dat layout written with the zlib stream deflated
// by several host threads.
//
// The reference's writer (include/MMSP.grid.cpp:1818-1909) runs one
// compress2(level 9) over the whole payload on one core, about 16 MB/s on
// near-random doubles: a 256^3 FP64 grid takes seconds to checkpoint while the
// GPU advances it 2000 steps in a tenth of a second. Here the payload is cut
// into chunks of whole nodes; each chunk is deflated on its own (raw deflate,
// level 9) with the 32 KiB of payload before it preset as its dictionary, and
// ends on a sync flush (an empty stored block, byte-aligned) except the last,
// which ends the stream. Concatenated behind the 2-byte zlib header and
// followed by the Adler-32 of the whole payload (combined from the chunks'),
// they form one zlib stream (RFC 1950/1951), so the file keeps the serial
// layout -- one block -- and every reader that uncompresses the block
// (grid::input, MMSP.grid.cpp:1306-1330; mmsp2png/mmsp2vti/mmsp2tsv) reads it.
// Only the compressed bytes differ from compress2's; the payload is the same.
//
// The chunks are written to the file in order as they complete. Workers wait
// for the node-major bytes they need through a caller-supplied function, so a
// download from the device can still be in flight while the first chunks are
// deflated.
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "dat.hpp"

namespace mmspx {

namespace {

constexpr size_t kWindow = 32768;  // deflate's window: the dictionary a chunk presets
constexpr int kSkipped = 1;        // chunk left undone after another one failed

struct Layout {
	int64_t nodes;
	int q;              // values per node
	size_t elem;        // bytes per value
	size_t node_bytes;  // payload bytes per node
	bool vector;        // vector<T> nodes carry an int32 count first
};

Layout layout_of(const mmspx_desc& d, size_t elem)
{
	Layout L;
	L.nodes = 1;
	for (int k = 0; k < d.dim; k++) L.nodes *= d.n[k];
	L.q = d.vector ? d.fields : 1;
	L.elem = elem;
	L.vector = d.vector != 0;
	L.node_bytes = L.vector ? sizeof(int32_t) + (size_t)L.q * elem : elem;
	return L;
}

// payload bytes of nodes [a, b) (MMSP.grid.cpp:1136-1157, MMSP.vector.cpp:100-105)
void payload(const Layout& L, const char* host, int64_t a, int64_t b, char* out)
{
	if (!L.vector) {
		std::memcpy(out, host + (size_t)a * L.elem, (size_t)(b - a) * L.elem);
		return;
	}
	const int32_t qq = L.q;
	const size_t vb = (size_t)L.q * L.elem;
	for (int64_t i = a; i < b; i++) {
		std::memcpy(out, &qq, sizeof(qq));
		std::memcpy(out + sizeof(qq), host + (size_t)i * vb, vb);
		out += L.node_bytes;
	}
}

struct Chunk {
	int64_t a = 0, b = 0;  // nodes
	std::vector<unsigned char> z;
	uLong adler = 1;
	int rc = 0;
	std::string err;
	bool done = false;
};

// Raw deflate of `in` at level 9 with `dict` preset; Z_SYNC_FLUSH, or
// Z_FINISH for the stream's last chunk.
int deflate_chunk(const unsigned char* dict, size_t dict_len, const unsigned char* in, size_t len, bool last,
                  std::vector<unsigned char>& out, std::string& err)
{
	z_stream zs;
	std::memset(&zs, 0, sizeof(zs));
	if (deflateInit2(&zs, 9, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
		err = "zlib deflateInit2 failed";
		return MMSPX_ENOMEM;
	}
	if (dict_len && deflateSetDictionary(&zs, dict, (uInt)dict_len) != Z_OK) {
		deflateEnd(&zs);
		err = "zlib deflateSetDictionary failed";
		return MMSPX_EIO;
	}
	const int flush = last ? Z_FINISH : Z_SYNC_FLUSH;
	out.resize(deflateBound(&zs, (uLong)len) + 64);
	zs.next_in = const_cast<unsigned char*>(in);
	zs.avail_in = (uInt)len;
	size_t have = 0;
	for (;;) {
		zs.next_out = out.data() + have;
		zs.avail_out = (uInt)(out.size() - have);
		const int r = deflate(&zs, flush);
		have = out.size() - zs.avail_out;
		if (r == Z_STREAM_ERROR) {
			deflateEnd(&zs);
			err = "zlib deflate failed";
			return MMSPX_EIO;
		}
		if (last ? r == Z_STREAM_END : (zs.avail_in == 0 && zs.avail_out != 0)) break;
		out.resize(out.size() * 2);
	}
	deflateEnd(&zs);
	out.resize(have);
	return 0;
}

std::string header_text(const mmspx_desc& d, size_t elem)
{
	// MMSP.grid.cpp:1370-1388; default stream precision (6 digits) for dx
	std::ostringstream h;
	const std::string t = elem == 8 ? "double" : "float";
	h << (d.vector ? "grid:vector:" + t : "grid:" + t) << '\n' << d.dim << '\n' << d.fields << '\n';
	for (int k = 0; k < d.dim; k++) h << d.g0[k] << " " << d.g0[k] + d.n[k] << '\n';
	for (int k = 0; k < d.dim; k++) h << d.dx[k] << '\n';
	return h.str();
}

bool put(FILE* f, const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }

}  // namespace

int dat_write_par(const char* filename, const mmspx_desc& d, const void* node_major, size_t elem, int compress,
                  int threads, const std::function<bool(size_t)>& wait_bytes)
{
	const Layout L = layout_of(d, elem);
	const char* host = (const char*)node_major;
	const uint64_t size_in_mem = (uint64_t)L.nodes * L.node_bytes;
	threads = std::max(1, threads);

	// Chunks of whole nodes: about eight per thread, between 256 KiB and 16 MiB.
	size_t target = (size_t)(size_in_mem / ((uint64_t)threads * 8) + 1);
	target = std::min<size_t>(std::max<size_t>(target, 256 * 1024), 16u << 20);
	const int64_t per = std::max<int64_t>(1, (int64_t)(target / L.node_bytes));
	std::vector<Chunk> chunks((size_t)((L.nodes + per - 1) / per));
	for (size_t c = 0; c < chunks.size(); c++) {
		chunks[c].a = (int64_t)c * per;
		chunks[c].b = std::min<int64_t>(L.nodes, chunks[c].a + per);
	}
	const int64_t dict_nodes = (int64_t)((kWindow + L.node_bytes - 1) / L.node_bytes);
	const size_t host_node_bytes = (size_t)L.q * L.elem;

	FILE* f = std::fopen(filename, "wb");
	if (!f) return fail(MMSPX_EIO, std::string("could not open ") + filename);

	std::mutex mu;
	std::condition_variable cv;
	std::atomic<size_t> next{0};
	std::atomic<bool> abort{false};

	auto work = [&]() {
		std::vector<char> buf;
		for (;;) {
			const size_t c = next.fetch_add(1);
			if (c >= chunks.size()) return;
			Chunk& ch = chunks[c];
			if (abort.load()) {
				// another chunk failed: mark this one so no one waits for it
				{
					std::lock_guard<std::mutex> g(mu);
					ch.rc = kSkipped;
					ch.done = true;
				}
				cv.notify_all();
				continue;
			}
			const int64_t a0 = std::max<int64_t>(0, ch.a - (compress ? dict_nodes : 0));
			bool ok = wait_bytes((size_t)ch.b * host_node_bytes);
			if (ok && compress) {
				const char* p;
				if (!L.vector) {
					p = host + (size_t)a0 * L.elem;  // scalar payload is the node-major array itself
				} else {
					buf.resize((size_t)(ch.b - a0) * L.node_bytes);
					payload(L, host, a0, ch.b, buf.data());
					p = buf.data();
				}
				const size_t pre = (size_t)(ch.a - a0) * L.node_bytes;
				const size_t dl = std::min(pre, kWindow);
				const unsigned char* in = (const unsigned char*)p + pre;
				const size_t len = (size_t)(ch.b - ch.a) * L.node_bytes;
				ch.rc = deflate_chunk(in - dl, dl, in, len, c + 1 == chunks.size(), ch.z, ch.err);
				ch.adler = adler32(1, in, (uInt)len);
			} else if (ok) {
				ch.z.resize((size_t)(ch.b - ch.a) * L.node_bytes);
				payload(L, host, ch.a, ch.b, (char*)ch.z.data());
			} else {
				ch.rc = MMSPX_EDEVICE;
				ch.err = "the state's download failed";
			}
			{
				std::lock_guard<std::mutex> g(mu);
				ch.done = true;
			}
			cv.notify_all();
			if (ch.rc) abort.store(true);
		}
	};
	std::vector<std::thread> pool;
	for (int t = 0; t < threads; t++) pool.emplace_back(work);

	// header and block table (MMSP.grid.cpp:1388-1396, 1830-1873); size_on_disk
	// is written once the chunks are in
	bool io = true;
	const std::string h = header_text(d, elem);
	io = io && put(f, h.data(), h.size());
	const int32_t nb = 1;
	io = io && put(f, &nb, 4);
	for (int k = 0; k < d.dim; k++) {
		const int32_t lim[2] = {d.g0[k], d.g0[k] + d.n[k]};
		io = io && put(f, lim, 8);
	}
	for (int k = 0; k < d.dim; k++) {
		const int32_t bc[2] = {d.b0[k], d.b1[k]};
		io = io && put(f, bc, 8);
	}
	io = io && put(f, &size_in_mem, 8);
	const long size_pos = std::ftell(f);
	uint64_t size_on_disk = 0;
	io = io && put(f, &size_on_disk, 8);
	if (compress) {
		const unsigned char zh[2] = {0x78, 0xDA};  // 32 KiB window, level 9 (RFC 1950 CMF/FLG)
		io = io && put(f, zh, 2);
		size_on_disk += 2;
	}
	uLong adler = 1;
	int rc = 0;
	std::string err;
	for (size_t c = 0; c < chunks.size() && !rc; c++) {
		Chunk& ch = chunks[c];
		{
			std::unique_lock<std::mutex> g(mu);
			cv.wait(g, [&] { return ch.done; });
		}
		if (ch.rc) {
			rc = ch.rc;
			err = ch.err;
			break;
		}
		io = io && put(f, ch.z.data(), ch.z.size());
		size_on_disk += ch.z.size();
		if (compress) adler = adler32_combine(adler, ch.adler, (z_off_t)((ch.b - ch.a) * (int64_t)L.node_bytes));
		std::vector<unsigned char>().swap(ch.z);
	}
	abort.store(abort.load() || rc != 0);
	for (auto& t : pool) t.join();
	if (rc) {
		// report the chunk that failed, not one skipped after it
		for (const Chunk& ch : chunks)
			if (ch.rc && ch.rc != kSkipped) {
				rc = ch.rc;
				err = ch.err;
				break;
			}
		std::fclose(f);
		std::remove(filename);
		return fail(rc, err);
	}
	if (compress) {
		const unsigned char tr[4] = {(unsigned char)(adler >> 24), (unsigned char)(adler >> 16),
		                             (unsigned char)(adler >> 8), (unsigned char)adler};
		io = io && put(f, tr, 4);
		size_on_disk += 4;
		if (size_on_disk == size_in_mem) {
			// a compressed block exactly the raw size would read back as raw
			io = io && std::fseek(f, size_pos + 8, SEEK_SET) == 0;
			std::vector<char> raw((size_t)size_in_mem);
			payload(L, host, 0, L.nodes, raw.data());
			io = io && put(f, raw.data(), raw.size());
		}
	}
	io = io && std::fseek(f, size_pos, SEEK_SET) == 0 && put(f, &size_on_disk, 8);
	io = (std::fclose(f) == 0) && io;
	if (!io) {
		std::remove(filename);
		return fail(MMSPX_EIO, std::string("write failed: ") + filename);
	}
	return 0;
}

}  // namespace mmspx
