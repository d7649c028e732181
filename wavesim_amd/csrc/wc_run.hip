// The run directory (format "wavesim-run", version 1) as waveio's RunWriter writes it
// (crates/waveio/src/run.rs:56-213): bed.f32 and frames.f32 as f32 little-endian, run.json
// last, printed as serde_json::to_string_pretty prints the RunHeader struct (run.rs:13-43).
//
// Frames taken on the GPU travel through two pinned host slots: the solver's stream copies a
// frame into a free slot, and the writer's thread writes the slot to frames.f32 while the
// caller goes on stepping. Frames given from the host are written by the caller's thread after
// the frames in flight, so the file keeps the order of the calls.

#include <errno.h>
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wavesim_run.h"
#include "wc_json.h"
#include "wc_run.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "f32-le is the host's own byte order");

namespace wc {
int api_fail(int code, const char* msg);  // wc_api.hip: sets wc_last_error
}

namespace {

using wc::json::Value;

constexpr const char* RUN_FORMAT = "wavesim-run";
constexpr uint64_t RUN_VERSION = 1;
constexpr const char* RUN_FILES[4] = {"run.json", "bed.f32", "frames.f32", "stats.f32"};
constexpr const char* LAYOUT =
    "per frame: one block per field, in `fields` order; each block row-major, y outer "
    "(increasing), x inner (increasing)";
constexpr int SLOTS = 2;

int fail(int code, const std::string& msg) { return wc::api_fail(code, msg.c_str()); }

// waveio's Error::Io / Error::Format messages name the path first.
int io_error(const std::string& path, int err) {
  return fail(WC_EIO, path + ": " + strerror(err));
}
int format_error(const std::string& path, const std::string& what) {
  return fail(WC_EFORMAT, path + ": " + what);
}

std::string join(const std::string& dir, const char* name) {
  return dir.empty() || dir.back() == '/' ? dir + name : dir + "/" + name;
}

// fs::create_dir_all
int create_dir_all(const std::string& dir) {
  std::string part;
  size_t q = 0;
  while (q <= dir.size()) {
    const size_t e = dir.find('/', q);
    part = dir.substr(0, e == std::string::npos ? dir.size() : e);
    q = e == std::string::npos ? dir.size() + 1 : e + 1;
    if (part.empty()) continue;
    if (mkdir(part.c_str(), 0777) != 0 && errno != EEXIST) return io_error(part, errno);
  }
  struct stat st;
  if (stat(dir.c_str(), &st) != 0) return io_error(dir, errno);
  if (!S_ISDIR(st.st_mode)) return io_error(dir, ENOTDIR);
  return WC_OK;
}

int write_file(const std::string& path, const void* data, size_t bytes) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return io_error(path, errno);
  const bool ok = bytes == 0 || fwrite(data, 1, bytes, f) == bytes;
  const int err = errno;
  if (fclose(f) != 0 || !ok) return io_error(path, ok ? errno : err);
  return WC_OK;
}

int parse_json(const char* text, const std::string& path, Value* out) {
  if (!text) {
    *out = Value::null();
    return WC_OK;
  }
  try {
    *out = wc::json::parse(text);
  } catch (const wc::json::ParseError& e) {
    return fail(WC_EFORMAT, path + ": invalid JSON: " + e.what());
  }
  return WC_OK;
}

struct StatsBlock {
  std::vector<std::string> fields;
  double from_s = 0.0, to_s = 0.0;
};

}  // namespace

struct wc_run_writer {
  std::string dir;
  // the header (run.rs:13-43)
  std::string name;
  uint64_t nx = 0, ny = 0;
  double dx = 0.0, dy = 0.0;
  std::vector<std::string> fields;
  std::vector<double> times;
  std::string crs, frame_note;
  double origin_easting = 0.0, origin_northing = 0.0, x_bearing_deg = 0.0;
  Value waves, source;
  bool has_stats = false;
  StatsBlock stats;
  std::vector<float> elevation;

  FILE* frames = nullptr;
  std::vector<char> frames_buffer;

  // frames from the GPU
  int device = -1;
  float* d_frame = nullptr;
  float* d_elevation = nullptr;
  unsigned long long* d_scalars = nullptr;
  float* slot[SLOTS] = {nullptr, nullptr};
  unsigned long long* scalars_host = nullptr;
  bool slot_busy[SLOTS] = {false, false};

  // the writer's thread
  std::thread thread;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int> queue;  // slots to write, in order
  bool stop = false;
  int io_errno = 0;  // the first failed write's errno

  uint64_t cells() const { return nx * ny; }
  std::string path(const char* file) const { return join(dir, file); }

  void start_thread() {
    if (thread.joinable()) return;
    thread = std::thread([this] { drain_loop(); });
  }

  void drain_loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv.wait(lk, [&] { return stop || !queue.empty(); });
      if (queue.empty()) return;  // stop, and nothing left
      const int s = queue.front();
      lk.unlock();
      const size_t bytes = fields.size() * cells() * sizeof(float);
      int err = 0;
      if (fwrite(slot[s], 1, bytes, frames) != bytes) err = errno ? errno : EIO;
      lk.lock();
      if (err && !io_errno) io_errno = err;
      queue.pop_front();
      slot_busy[s] = false;
      cv.notify_all();
    }
  }

  // Waits until every frame in flight is in the file's buffer; the first write error, if any.
  int wait_idle() {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return queue.empty(); });
    return io_errno ? io_error(path("frames.f32"), io_errno) : WC_OK;
  }

  void stop_thread() {
    if (!thread.joinable()) return;
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    thread.join();
  }

  void release() {
    stop_thread();
    if (frames) fclose(frames);
    frames = nullptr;
    if (device >= 0) {
      (void)hipSetDevice(device);
      (void)hipFree(d_frame);
      (void)hipFree(d_elevation);
      (void)hipFree(d_scalars);
      for (float* p : slot) (void)hipHostFree(p);
      (void)hipHostFree(scalars_host);
      device = -1;
    }
  }

  std::string header_json() const {
    wc::json::Fields h(0);
    h.value("format", Value::string(RUN_FORMAT));
    h.value("version", Value::uint(RUN_VERSION));
    h.value("name", Value::string(name));
    h.value("nx", Value::uint(nx));
    h.value("ny", Value::uint(ny));
    h.value("dx", Value::number(dx));
    h.value("dy", Value::number(dy));
    h.value("bed", Value::string("bed.f32"));
    h.value("frames", Value::string("frames.f32"));
    Value fl;
    fl.kind = Value::Array;
    for (const auto& f : fields) fl.a.push_back(Value::string(f));
    h.value("fields", fl);
    h.value("dtype", Value::string("f32-le"));
    h.value("layout", Value::string(LAYOUT));
    h.value("frame_count", Value::uint(times.size()));
    Value tm;
    tm.kind = Value::Array;
    for (double t : times) tm.a.push_back(Value::number(t));
    h.value("times", tm);
    wc::json::Fields fr(1);
    fr.value("crs", Value::string(crs));
    fr.value("origin_easting", Value::number(origin_easting));
    fr.value("origin_northing", Value::number(origin_northing));
    fr.value("x_bearing_deg", Value::number(x_bearing_deg));
    fr.value("note", Value::string(frame_note));
    h.raw("frame", fr.close());
    h.value("waves", waves);
    h.value("source", source);
    if (has_stats) {  // skip_serializing_if = "Option::is_none"
      wc::json::Fields st(1);
      st.value("file", Value::string("stats.f32"));
      Value sf;
      sf.kind = Value::Array;
      for (const auto& f : stats.fields) sf.a.push_back(Value::string(f));
      st.value("fields", sf);
      st.value("from_s", Value::number(stats.from_s));
      st.value("to_s", Value::number(stats.to_s));
      h.raw("stats", st.close());
    }
    return h.close();
  }
};

namespace wc::run {

int begin_device_frame(wc_run_writer* w, int device, long long nx, long long ny,
                       DeviceFrame* f) {
  if ((uint64_t)nx != w->nx || (uint64_t)ny != w->ny)
    return format_error(w->path("frames.f32"), "the solver's grid is " + std::to_string(nx) +
                                                    " x " + std::to_string(ny) + ", the run's " +
                                                    std::to_string(w->nx) + " x " +
                                                    std::to_string(w->ny));
  int eta = -1, brk = -1;
  for (size_t q = 0; q < w->fields.size(); ++q) {
    if (w->fields[q] == "eta") eta = (int)q;
    if (w->fields[q] == "breaking") brk = (int)q;
  }
  if (w->fields.size() != 2 || eta < 0 || brk < 0)
    return format_error(w->path("frames.f32"),
                        "a frame from the GPU holds the fields \"eta\" and \"breaking\"");
  if (w->device >= 0 && w->device != device)
    return fail(WC_EINVAL, "a run's frames come from one device");
  const size_t n = w->cells();
  if (w->device < 0) {
    auto hip = [&](hipError_t e) {
      return e == hipSuccess ? WC_OK
                             : fail(e == hipErrorOutOfMemory ? WC_ENOMEM : WC_EDEVICE,
                                    std::string("run frame buffers: ") + hipGetErrorString(e));
    };
    w->device = device;
    int rc;
    if ((rc = hip(hipMalloc(&w->d_frame, 2 * n * sizeof(float)))) ||
        (rc = hip(hipMalloc(&w->d_elevation, n * sizeof(float)))) ||
        (rc = hip(hipMalloc(&w->d_scalars, 2 * sizeof(unsigned long long)))) ||
        (rc = hip(hipHostMalloc(&w->slot[0], 2 * n * sizeof(float), hipHostMallocDefault))) ||
        (rc = hip(hipHostMalloc(&w->slot[1], 2 * n * sizeof(float), hipHostMallocDefault))) ||
        (rc = hip(hipHostMalloc(&w->scalars_host, 2 * sizeof(unsigned long long),
                                hipHostMallocDefault))) ||
        (rc = hip(hipMemcpy(w->d_elevation, w->elevation.data(), n * sizeof(float),
                            hipMemcpyHostToDevice))))
      return rc;
    w->start_thread();
  }
  int s = -1;
  {
    std::unique_lock<std::mutex> lk(w->mu);
    w->cv.wait(lk, [&] { return !w->slot_busy[0] || !w->slot_busy[1] || w->io_errno; });
    if (w->io_errno) return io_error(w->path("frames.f32"), w->io_errno);
    s = w->slot_busy[0] ? 1 : 0;
    w->slot_busy[s] = true;
  }
  f->slot = s;
  f->cells = n;
  f->eta = eta;
  f->breaking = brk;
  f->device = w->d_frame;
  f->pinned = w->slot[s];
  f->elevation = w->d_elevation;
  f->scalars = w->d_scalars;
  f->scalars_host = w->scalars_host;
  return WC_OK;
}

int commit_frame(wc_run_writer* w, const DeviceFrame& f, double time) {
  {
    std::lock_guard<std::mutex> lk(w->mu);
    w->queue.push_back(f.slot);
  }
  w->cv.notify_all();
  w->times.push_back(time);
  return WC_OK;
}

void abandon_frame(wc_run_writer* w, const DeviceFrame& f) {
  {
    std::lock_guard<std::mutex> lk(w->mu);
    w->slot_busy[f.slot] = false;
  }
  w->cv.notify_all();
}

void note(wc_run_writer* w, const char* key, Value value) {
  if (w->waves.kind != Value::Object) w->waves = Value::object();
  w->waves.o[key] = std::move(value);
}

}  // namespace wc::run

extern "C" {

int wc_remove_run(const char* dir) {
  if (!dir) return fail(WC_EINVAL, "null directory");
  for (const char* name : RUN_FILES) {
    const std::string p = join(dir, name);
    if (unlink(p.c_str()) != 0 && errno != ENOENT) return io_error(p, errno);
  }
  return WC_OK;
}

int wc_run_create(const char* dir, const char* name, uint64_t nx, uint64_t ny, double dx,
                  double dy, const float* elevation, const wc_frame_t* frame,
                  const char* source_json, const char* const* fields, uint64_t n_fields,
                  const char* waves_json, wc_run_writer** out) {
  if (!dir || !out || (!fields && n_fields) || (!elevation && nx * ny))
    return fail(WC_EINVAL, "null argument");
  *out = nullptr;
  bool has_eta = false;
  for (uint64_t q = 0; q < n_fields; ++q) {
    if (!fields[q]) return fail(WC_EINVAL, "null field name");
    has_eta = has_eta || strcmp(fields[q], "eta") == 0;
  }
  if (!has_eta) return format_error(dir, "a run must have an \"eta\" field");
  auto w = std::make_unique<wc_run_writer>();
  w->dir = dir;
  int rc;
  if ((rc = parse_json(waves_json, w->path("run.json"), &w->waves))) return rc;
  if ((rc = parse_json(source_json, w->path("run.json"), &w->source))) return rc;
  if ((rc = create_dir_all(w->dir))) return rc;
  if ((rc = wc_remove_run(dir))) return rc;
  w->name = name ? name : "";
  w->nx = nx;
  w->ny = ny;
  w->dx = dx;
  w->dy = dy;
  for (uint64_t q = 0; q < n_fields; ++q) w->fields.push_back(fields[q]);
  if (frame) {
    w->crs = frame->crs ? frame->crs : "";
    w->origin_easting = frame->origin_easting;
    w->origin_northing = frame->origin_northing;
    w->x_bearing_deg = frame->x_bearing_deg;
    w->frame_note = frame->note ? frame->note : "";
  } else {
    w->crs = "local";  // Bed::header_for's frame (bed.rs:152-158)
  }
  w->elevation.assign(elevation, elevation + nx * ny);
  if ((rc = write_file(w->path("bed.f32"), w->elevation.data(),
                       w->elevation.size() * sizeof(float))))
    return rc;
  const std::string fp = w->path("frames.f32");
  w->frames = fopen(fp.c_str(), "wb");
  if (!w->frames) return io_error(fp, errno);
  w->frames_buffer.resize(1 << 20);
  setvbuf(w->frames, w->frames_buffer.data(), _IOFBF, w->frames_buffer.size());
  *out = w.release();
  return WC_OK;
}

int wc_run_note(wc_run_writer* w, const char* key, const char* value_json) {
  if (!w || !key) return fail(WC_EINVAL, "null argument");
  Value v;
  int rc;
  if ((rc = parse_json(value_json, w->path("run.json"), &v))) return rc;
  wc::run::note(w, key, std::move(v));
  return WC_OK;
}

int wc_run_write_frame(wc_run_writer* w, double time, const float* const* fields,
                       uint64_t n_fields, const uint64_t* lens) {
  if (!w || (!fields && n_fields)) return fail(WC_EINVAL, "null argument");
  const std::string path = w->path("frames.f32");
  const uint64_t want = w->cells();
  if (n_fields != w->fields.size())
    return format_error(path, "frame has " + std::to_string(n_fields) + " fields, expected " +
                                  std::to_string(w->fields.size()));
  for (uint64_t q = 0; q < n_fields; ++q) {
    const uint64_t len = lens ? lens[q] : want;
    if (len != want)
      return format_error(path, "a field has " + std::to_string(len) + " values, expected " +
                                    std::to_string(want));
    if (!fields[q] && want) return fail(WC_EINVAL, "null field");
  }
  int rc;
  if ((rc = w->wait_idle())) return rc;
  for (uint64_t q = 0; q < n_fields; ++q)
    if (want && fwrite(fields[q], sizeof(float), want, w->frames) != want)
      return io_error(path, errno);
  w->times.push_back(time);
  return WC_OK;
}

int wc_run_write_stats(wc_run_writer* w, const char* const* fields, const float* const* blocks,
                       uint64_t n_fields, const uint64_t* lens, double from, double to) {
  if (!w || ((!fields || !blocks) && n_fields)) return fail(WC_EINVAL, "null argument");
  const std::string path = w->path("stats.f32");
  const uint64_t want = w->cells();
  for (uint64_t q = 0; q < n_fields; ++q)
    if ((lens ? lens[q] : want) != want || !fields[q] || (!blocks[q] && want))
      return format_error(path, "stats must be " + std::to_string(n_fields) + " blocks of " +
                                    std::to_string(want) + " values");
  std::vector<float> all;
  all.reserve(n_fields * want);
  for (uint64_t q = 0; q < n_fields; ++q) all.insert(all.end(), blocks[q], blocks[q] + want);
  int rc;
  if ((rc = write_file(path, all.data(), all.size() * sizeof(float)))) return rc;
  w->has_stats = true;
  w->stats.fields.clear();
  for (uint64_t q = 0; q < n_fields; ++q) w->stats.fields.push_back(fields[q]);
  w->stats.from_s = from;
  w->stats.to_s = to;
  return WC_OK;
}

int wc_run_header(const wc_run_writer* w, uint64_t* frame_count, char* json_out, uint64_t cap,
                  uint64_t* needed) {
  if (!w) return fail(WC_EINVAL, "null writer");
  if (frame_count) *frame_count = w->times.size();
  const std::string text = w->header_json() + "\n";
  if (needed) *needed = text.size() + 1;
  if (json_out && cap) {
    const size_t n = std::min<size_t>(cap - 1, text.size());
    memcpy(json_out, text.data(), n);
    json_out[n] = 0;
  }
  return WC_OK;
}

int wc_run_finish(wc_run_writer* w, char* path_out, uint64_t path_cap) {
  if (!w) return fail(WC_EINVAL, "null writer");
  std::unique_ptr<wc_run_writer> own(w);
  const std::string fp = w->path("frames.f32");
  int rc = w->wait_idle();
  w->stop_thread();
  if (rc) {
    w->release();
    return rc;
  }
  const bool flushed = fflush(w->frames) == 0;
  const int err = errno;
  const bool closed = fclose(w->frames) == 0;
  w->frames = nullptr;
  if (!flushed || !closed) {
    w->release();
    return io_error(fp, flushed ? errno : err);
  }
  w->release();
  const std::string path = w->path("run.json");
  const std::string text = w->header_json() + "\n";
  if ((rc = write_file(path, text.data(), text.size()))) return rc;
  if (path_out && path_cap) {
    const size_t n = std::min<size_t>(path_cap - 1, path.size());
    memcpy(path_out, path.data(), n);
    path_out[n] = 0;
  }
  return WC_OK;
}

void wc_run_free(wc_run_writer* w) {
  if (!w) return;
  (void)w->wait_idle();
  w->release();
  delete w;
}

}  // extern "C"
