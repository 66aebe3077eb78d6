/* edenhip_output.h -- EDEN's recorded-output text formats for the rows and events that
 * edenhip_read_trace / edenhip_read_events return.
 *
 * Replaces the reference's TimeseriesWriter / EventWriter (eden/Eden.cpp:10843-10940) with
 * the engine loop's FixedWidthNumberPrinter(16, '\t', 0) (eden/Eden.cpp:144-181,
 * 10570-10571): every number fills a 16-character field, "%+16.8g", with no delimiter of its
 * own. Timestamps are the engine time (us) times 1e-6 (seconds,
 * GetTimestampAscii); trajectory values are the recorded float times each column's scale
 * factor (native unit to the output unit: 1e-3 for mV to V in a classic OutputFile).
 *
 *   trajectory line:  time ' ' value_0 ' ' value_1 ...            (then '\n')
 *   events, TIME_ID:  one line per event: time ' ' id
 *   events, ID_TIME:  one line per event: id ' ' time (time without padding)
 *   events, ascii_v0: one line per row with events: time ' ' id ' ' id ...
 * Event lines follow the rows in ascending order and, within a row, the columns in
 * ascending order. */
#ifndef EDENHIP_OUTPUT_H
#define EDENHIP_OUTPUT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum edenhip_event_format { EDENHIP_EVENTS_TIME_ID = 0, EDENHIP_EVENTS_ID_TIME = 1, EDENHIP_EVENTS_ASCII_V0 = 2 };

/* Append rows [0, n_rows) of trace (n_rows x n_cols floats, row-major) to the file at
 * `path` (created or truncated when `append` is 0). Row r is stamped with time_us[r];
 * column c is scaled by scale[c]. Returns 0, or -1 for bad arguments, -2 for an I/O error. */
int edenhip_write_trajectory(const char *path, int append, const double *time_us, const float *trace, int64_t n_rows,
                             int64_t n_cols, const double *scale);

/* Write events (row, column) sorted by row then column, row r stamped with time_us[r] and
 * column c printed as ids[c], in the given edenhip_event_format. */
int edenhip_write_events(const char *path, int append, int32_t format, const double *time_us, const int64_t *ev_row,
                         const int64_t *ev_col, int64_t n_events, const int64_t *ids, int64_t n_ids);

/* The fixed-width text of one number as the writers print it ("%+16.8g": 16 characters plus
 * the terminator, so buf_len >= 17): for callers building their own lines. */
int edenhip_format_number(double value, char *buf, int32_t buf_len);

#ifdef __cplusplus
}
#endif
#endif /* EDENHIP_OUTPUT_H */
