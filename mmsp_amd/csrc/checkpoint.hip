// checkpoint.hip -- .dat checkpoints written while the state keeps stepping.
//
// The reference's driver (MMSP.main.hpp:278-350) alternates update(grid,
// increment) with grid.output(filename), and output() deflates the whole grid
// on one core at zlib level 9 (MMSP.grid.cpp:1358-1402, 1818-1909). Here a
// checkpoint takes a device-side snapshot of the state on the state's own
// stream (a copy, or the field-major -> node-major transpose for vector
// grids, into a stream-ordered allocation) and returns. Steps enqueued after
// it run against the state's own buffers, so the simulation goes on while a
// host thread downloads the snapshot slab by slab on a stream of its own and
// a pool of host threads deflates it in chunks into the serial .dat layout
// (dat_par.cpp). One thread writes compress2's exact bytes (dat.cpp).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "core.hpp"

struct mmspx_checkpoint {
	std::thread th;
	int rc = 0;
	std::string err;
	double seconds = 0;
	mmspx_desc desc{};
	size_t elem = 0, bytes = 0;
	int device = 0;
	std::string filename;
	int compress = 1, threads = 1;
	void* snap = nullptr;
	hipEvent_t ev = nullptr;
	std::chrono::steady_clock::time_point t0;
};

namespace mmspx {

namespace {

constexpr size_t kSlabBytes = 16u << 20;  // download granularity

int host_threads(int threads)
{
	if (threads > 0) return std::min(threads, 256);
	const unsigned h = std::thread::hardware_concurrency();
	return (int)std::max(1u, std::min(h, 64u));
}

// Downloads the snapshot while the writer deflates what has arrived.
void run_checkpoint(mmspx_checkpoint* c)
{
	std::unique_ptr<char[]> host(new (std::nothrow) char[c->bytes]);
	std::mutex mu;
	std::condition_variable cv;
	size_t ready = 0;
	bool failed = false;
	std::string derr;

	auto download = [&]() {
		hipStream_t cs = nullptr;
		hipError_t e = hipSetDevice(c->device);
		if (e == hipSuccess) e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
		if (e == hipSuccess) e = hipStreamWaitEvent(cs, c->ev, 0);
		for (size_t off = 0; e == hipSuccess && off < c->bytes; off += kSlabBytes) {
			const size_t len = std::min(kSlabBytes, c->bytes - off);
			e = hipMemcpyAsync(host.get() + off, (const char*)c->snap + off, len, hipMemcpyDeviceToHost, cs);
			if (e == hipSuccess) e = hipStreamSynchronize(cs);
			if (e == hipSuccess) {
				std::lock_guard<std::mutex> g(mu);
				ready = off + len;
			}
			cv.notify_all();
		}
		if (cs) {
			const hipError_t f = hipFreeAsync(c->snap, cs);
			if (e == hipSuccess) e = f;
			const hipError_t s = hipStreamSynchronize(cs);
			if (e == hipSuccess) e = s;
			(void)hipStreamDestroy(cs);
		} else {
			(void)hipEventSynchronize(c->ev);
			(void)hipFree(c->snap);
		}
		c->snap = nullptr;
		std::lock_guard<std::mutex> g(mu);
		if (e != hipSuccess) {
			failed = true;
			derr = std::string("checkpoint download: ") + hipGetErrorString(e);
		}
		ready = c->bytes;
		cv.notify_all();
	};
	auto wait_bytes = [&](size_t n) {
		std::unique_lock<std::mutex> g(mu);
		cv.wait(g, [&] { return ready >= n || failed; });
		return !failed;
	};

	if (!host) {
		// nothing can be written; still release the snapshot once it is made
		(void)hipSetDevice(c->device);
		(void)hipEventSynchronize(c->ev);
		(void)hipFree(c->snap);
		c->snap = nullptr;
		c->rc = MMSPX_ENOMEM;
		c->err = "checkpoint: host allocation failed";
	} else {
		std::thread dl(download);
		if (c->threads == 1 && c->compress) {
			// one thread: compress2 over the whole payload, the reference's exact bytes
			if (!wait_bytes(c->bytes)) {
				c->rc = MMSPX_EDEVICE;
			} else {
				c->rc = dat_write(c->filename.c_str(), c->desc, host.get(), c->elem, 1);
			}
		} else {
			c->rc = dat_write_par(c->filename.c_str(), c->desc, host.get(), c->elem, c->compress, c->threads,
			                      wait_bytes);
		}
		if (c->rc) c->err = mmspx_last_error();
		dl.join();
		if (failed) {
			c->rc = MMSPX_EDEVICE;
			c->err = derr;
			std::remove(c->filename.c_str());
		}
	}
	(void)hipEventDestroy(c->ev);
	c->ev = nullptr;
	c->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - c->t0).count();
}

}  // namespace

}  // namespace mmspx

using namespace mmspx;

extern "C" {

int mmspx_checkpoint_begin(mmspx_state* s, const char* filename, int compress, int threads, mmspx_checkpoint** out)
{
	if (!out) return fail(MMSPX_EINVAL, "null output pointer");
	*out = nullptr;
	if (!s || !filename) return fail(MMSPX_EINVAL, "null argument");
	if (compress != 0 && compress != 1) return fail(MMSPX_EINVAL, "compress must be 0 or 1");
	if (threads < 0) return fail(MMSPX_EINVAL, "threads must be >= 0 (0: one per host core)");
	std::unique_ptr<mmspx_checkpoint> c(new (std::nothrow) mmspx_checkpoint());
	if (!c) return fail(MMSPX_ENOMEM, "host allocation failed");
	c->t0 = std::chrono::steady_clock::now();
	c->desc = s->desc;
	c->elem = s->elem;
	c->bytes = s->bytes;
	c->device = s->device;
	c->filename = filename;
	c->compress = compress;
	c->threads = host_threads(threads);
	MMSPX_HIP(hipSetDevice(s->device));
	MMSPX_HIP(hipMallocAsync(&c->snap, s->bytes, s->stream));
	int rc = 0;
	if (s->geom.fields == 1) {
		if (hipMemcpyAsync(c->snap, s->cur, s->bytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
			rc = fail(MMSPX_EDEVICE, "checkpoint snapshot copy failed");
	} else {
		rc = launch_transpose(s->cur, c->snap, s->geom.nodes, s->geom.fields, s->elem, false, s->stream);
	}
	if (!rc && hipEventCreateWithFlags(&c->ev, hipEventDisableTiming) != hipSuccess)
		rc = fail(MMSPX_EDEVICE, "hipEventCreate failed");
	if (!rc && hipEventRecord(c->ev, s->stream) != hipSuccess) rc = fail(MMSPX_EDEVICE, "hipEventRecord failed");
	if (rc) {
		(void)hipFreeAsync(c->snap, s->stream);
		if (c->ev) (void)hipEventDestroy(c->ev);
		return rc;
	}
	mmspx_checkpoint* p = c.release();
	try {
		p->th = std::thread(run_checkpoint, p);
	} catch (...) {
		(void)hipStreamSynchronize(s->stream);
		(void)hipFree(p->snap);
		(void)hipEventDestroy(p->ev);
		delete p;
		return fail(MMSPX_ENOMEM, "could not start the checkpoint thread");
	}
	*out = p;
	return 0;
}

int mmspx_checkpoint_wait(mmspx_checkpoint* c, double* seconds)
{
	if (!c) return fail(MMSPX_EINVAL, "null checkpoint");
	if (c->th.joinable()) c->th.join();
	const int rc = c->rc;
	const std::string err = c->err;
	if (seconds) *seconds = c->seconds;
	delete c;
	if (rc) return fail(rc, err);
	return 0;
}

int mmspx_write_dat_host(const mmspx_desc* d, const void* node_major, const char* filename, int compress, int threads)
{
	if (!d || !node_major || !filename) return fail(MMSPX_EINVAL, "null argument");
	int rc = validate_desc(d);
	if (rc) return rc;
	if (compress != 0 && compress != 1) return fail(MMSPX_EINVAL, "compress must be 0 or 1");
	if (threads < 0) return fail(MMSPX_EINVAL, "threads must be >= 0 (0: one per host core)");
	const size_t elem = d->precision == MMSPX_F64 ? 8 : 4;
	const int t = host_threads(threads);
	if (t == 1) return dat_write(filename, *d, node_major, elem, compress);
	return dat_write_par(filename, *d, node_major, elem, compress, t, [](size_t) { return true; });
}

}  // extern "C"
