/* wavesim_run — the run directory the viewer reads (format "wavesim-run", version 1), written
 * as waveio's RunWriter writes it (crates/waveio/src/run.rs:72-213; docs/data-format.md
 * "Run"), with frames and statistics taken from a wc_solver / wc_stats on the GPU.
 *
 * A run directory holds run.json (the header), bed.f32 (the bed, f32 little-endian, row-major
 * with y outer), frames.f32 (frame_count frames of every field, back to back: field f of frame
 * k starts at value (k * n_fields + f) * nx * ny) and, when statistics were written,
 * stats.f32 (one block of nx * ny values per statistics field). run.json is written last, by
 * wc_run_finish, so a run that stopped part-way has frames but no header.
 *
 * The JSON values a caller hands over (the bed's source, the run's waves, notes) are JSON
 * text; run.json prints them the way the reference's serde_json does (object keys sorted,
 * floats in their shortest round-trip digits), so the same run gives the same bytes.
 *
 * Frames taken from the GPU (wc_run_frame_from) are computed by one kernel, copied to pinned
 * host memory and written to disk by the writer's own thread while the caller goes on
 * stepping; every other call is synchronous. A writer is used from one thread at a time.
 *
 * Errors follow wavecore_hip.h: WC_OK or a code, with the message in wc_last_error().
 * WC_EIO is waveio's Error::Io (a file could not be written or removed), WC_EFORMAT its
 * Error::Format (the data does not fit the run) and Error::Json (a JSON value that does not
 * parse).
 */
#ifndef WAVESIM_RUN_H
#define WAVESIM_RUN_H

#include <stdint.h>

#include "wavecore_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define WC_EIO 4
#define WC_EFORMAT 5

typedef struct wc_run_writer wc_run_writer;

/* Frame (crates/waveio/src/bed.rs:14-24): where the grid sits on the Earth. NULL strings are
 * empty. */
typedef struct {
  const char* crs;
  double origin_easting, origin_northing, x_bearing_deg;
  const char* note;
} wc_frame_t;

/* remove_run (run.rs:56-68): removes run.json, bed.f32, frames.f32 and stats.f32 from dir, if
 * they exist, and nothing else. A missing directory or file is not an error. */
int wc_remove_run(const char* dir);

/* RunWriter::create (run.rs:80-124) for a bed (crates/waveio/src/bed.rs:26-60) of nx * ny
 * cells of dx * dy metres named `name`, with `elevation` its nx * ny f32 values, `frame` its
 * frame and `source_json` its source (NULL: null). `fields` names what each frame holds, in
 * order, and must include "eta"; `waves_json` (NULL: null) records the run's settings.
 * Creates dir, removes a run already there, writes bed.f32 and opens frames.f32. */
int wc_run_create(const char* dir, const char* name, uint64_t nx, uint64_t ny, double dx,
                  double dy, const float* elevation, const wc_frame_t* frame,
                  const char* source_json, const char* const* fields, uint64_t n_fields,
                  const char* waves_json, wc_run_writer** out);

/* RunWriter::note (run.rs:126-132): waves[key] = value (JSON text), making waves an object
 * first if it is not one. */
int wc_run_note(wc_run_writer* w, const char* key, const char* value_json);

/* RunWriter::write_frame (run.rs:134-163): appends one frame taken at `time`: n_fields host
 * arrays, field q holding lens[q] values (lens NULL: nx * ny each), in the run's field order.
 * Refused (WC_EFORMAT) when the count of fields or of values does not fit. */
int wc_run_write_frame(wc_run_writer* w, double time, const float* const* fields,
                       uint64_t n_fields, const uint64_t* lens);

/* RunWriter::write_stats (run.rs:165-196): writes stats.f32, n_fields blocks (block q of
 * lens[q] values, lens NULL: nx * ny each) named `fields`, recorded from `from` to `to`
 * seconds, and records them in the header. */
int wc_run_write_stats(wc_run_writer* w, const char* const* fields, const float* const* blocks,
                       uint64_t n_fields, const uint64_t* lens, double from, double to);

/* RunWriter::finish (run.rs:198-212): waits for the frames in flight, flushes frames.f32 and
 * writes run.json. Copies the header's path into path_out (capacity path_cap bytes, may be 0)
 * and frees the writer whatever the outcome. */
int wc_run_finish(wc_run_writer* w, char* path_out, uint64_t path_cap);

/* Drops a writer without finishing it: the frames written so far stay, with no run.json (a
 * run that stopped part-way). */
void wc_run_free(wc_run_writer* w);

/* The header's frame count and the JSON text run.json would hold now (capacity cap bytes; the
 * length it needs, with its terminating NUL, is returned in *needed). */
int wc_run_header(const wc_run_writer* w, uint64_t* frame_count, char* json_out, uint64_t cap,
                  uint64_t* needed);

/* Stage::frame_if_due's frame (crates/wavesim/src/lib.rs:1348-1380) from the solver's current
 * state: eta = (h + elevation) as f32 with the run's f32 bed elevation, breaking = the
 * solver's breaking indicator (wc_breaking_indicator) as f32. *rise is the largest eta - tide
 * over the cells deeper than 1e-3 m (from 0). When rise is not finite or above 100 m, or an
 * eta is not finite, the frame is not written and *status is WC_STAGE_UNSTABLE (the
 * reference's Error::Unstable); otherwise it joins the frames at the solver's time and
 * *status is WC_STAGE_OK. The run's fields must be "eta" and "breaking", in either order.
 * The disk write happens on the writer's thread; its errors come back from a later call. */
int wc_run_frame_from(wc_run_writer* w, wc_solver* s, double tide, double* rise,
                      int32_t* status);

/* Stage::finish's statistics (lib.rs:1436-1509, without the break line): after closing the
 * wave in progress if it is due (wc_stats_finish), writes break_fraction, eta_max, eta_min
 * and, once a whole wave has been recorded, wave_height and, when the first harmonic can be
 * fitted, harmonic_re and harmonic_im, from `from` to the solver's time, and notes
 * waves_recorded. Writes nothing when no time was recorded. */
int wc_run_stats_from(wc_run_writer* w, wc_stats* st);

#ifdef __cplusplus
}
#endif

#endif
