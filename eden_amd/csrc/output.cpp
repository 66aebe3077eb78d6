// output.cpp -- EDEN's trajectory and event text formats (include/edenhip_output.h).
//
// Number format: the engine loop's printer is FixedWidthNumberPrinter(16, '\t', 0)
// (eden/Eden.cpp:10570-10571): a 16-character column with a zero-length delimiter, so the
// number takes all 16 characters, "%+16.8g" (8 = 16 - 3 - 5 significant digits,
// eden/Eden.cpp:144-181), and nothing follows it. Columns of a line are joined by one ' '
// (eden/Eden.cpp:10856-10870). Timestamps: engine time (us) times 1e-6 (GetTimestampAscii,
// eden/Eden.cpp:184-189). Lines end with '\n' (LineWriter_File::PutFrameLines,
// eden/Eden.cpp:10633-10640).
#include <stdint.h>

#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/edenhip_output.h"

namespace {

constexpr double US_TO_S = 1e-6;

// 16 characters: the number right-aligned in a 16-character field, no delimiter. The bound
// of 17 bytes (16 characters and the terminator) is the reference's snprintf bound; "%+.8g"
// never needs more than 15 characters, so nothing is cut.
inline void fixed16(double v, char *out) { std::snprintf(out, 17, "%+16.8g", v); }

struct File {
    FILE *f = nullptr;
    File(const char *path, int append) { f = std::fopen(path, append ? "ab" : "wb"); }
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

int edenhip_format_number(double value, char *buf, int32_t buf_len) {
    if (!buf || buf_len < 17) return -1;
    fixed16(value, buf);
    return 0;
}

int edenhip_write_trajectory(const char *path, int append, const double *time_us, const float *trace, int64_t n_rows,
                             int64_t n_cols, const double *scale) {
    if (!path || n_rows < 0 || n_cols < 0 || (n_rows > 0 && (!time_us || (n_cols > 0 && (!trace || !scale))))) return -1;
    File out(path, append);
    if (!out.f) return -2;
    // Rows are formatted in blocks by worker threads and written in row order: the text is
    // the same as one thread's, the formatting (most of the cost) runs in parallel.
    constexpr int64_t BLOCK = 64;
    const int64_t n_blocks = (n_rows + BLOCK - 1) / BLOCK;
    const int nthr = (int)std::max<int64_t>(1, std::min<int64_t>(n_blocks, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    auto format_block = [&](int64_t blk, std::string &line) {
        char num[32];
        line.clear();
        for (int64_t r = blk * BLOCK; r < std::min(n_rows, (blk + 1) * BLOCK); r++) {
            fixed16(time_us[r] * US_TO_S, num);
            line += num;
            for (int64_t c = 0; c < n_cols; c++) {
                line += ' ';
                fixed16((double)trace[r * n_cols + c] * scale[c], num);
                line += num;
            }
            line += '\n';
        }
    };
    if (nthr <= 1) {
        std::string line;
        for (int64_t b = 0; b < n_blocks; b++) {
            format_block(b, line);
            if (std::fwrite(line.data(), 1, line.size(), out.f) != line.size()) return -2;
        }
        return std::ferror(out.f) ? -2 : 0;
    }
    // a ring of formatted blocks: workers take block numbers in order; the writer (this
    // thread) writes block b once it is ready, then frees its slot for block b + ring
    const int64_t ring = 4 * nthr;
    std::vector<std::string> bufs((size_t)ring);
    std::vector<int> ready((size_t)ring, 0);
    std::mutex mu;
    std::condition_variable cv;
    int64_t next = 0, written = 0;
    bool failed = false;
    auto worker = [&]() {
        std::string tmp;
        for (;;) {
            int64_t b;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return failed || next >= n_blocks || next < written + ring; });
                if (failed || next >= n_blocks) return;
                b = next++;
            }
            format_block(b, tmp);
            {
                std::lock_guard<std::mutex> lk(mu);
                bufs[(size_t)(b % ring)].swap(tmp);
                ready[(size_t)(b % ring)] = 1;
            }
            cv.notify_all();
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nthr; t++) pool.emplace_back(worker);
    int rc = 0;
    for (int64_t b = 0; b < n_blocks; b++) {
        std::string blk;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return ready[(size_t)(b % ring)] != 0; });
            blk.swap(bufs[(size_t)(b % ring)]);
            ready[(size_t)(b % ring)] = 0;
        }
        if (rc == 0 && std::fwrite(blk.data(), 1, blk.size(), out.f) != blk.size()) rc = -2;
        {
            std::lock_guard<std::mutex> lk(mu);
            written = b + 1;
            if (rc != 0) failed = true;
        }
        cv.notify_all();
        if (rc != 0) break;
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        failed = failed || rc != 0;
    }
    cv.notify_all();
    for (auto &t : pool) t.join();
    if (rc != 0) return rc;
    return std::ferror(out.f) ? -2 : 0;
}

int edenhip_write_events(const char *path, int append, int32_t format, const double *time_us, const int64_t *ev_row,
                         const int64_t *ev_col, int64_t n_events, const int64_t *ids, int64_t n_ids) {
    if (!path || n_events < 0 || (n_events > 0 && (!time_us || !ev_row || !ev_col || !ids))) return -1;
    if (format != EDENHIP_EVENTS_TIME_ID && format != EDENHIP_EVENTS_ID_TIME && format != EDENHIP_EVENTS_ASCII_V0) return -1;
    for (int64_t e = 0; e < n_events; e++) {
        if (ev_col[e] < 0 || ev_col[e] >= n_ids || ev_row[e] < 0) return -1;
        if (e > 0 && (ev_row[e] < ev_row[e - 1] || (ev_row[e] == ev_row[e - 1] && ev_col[e] <= ev_col[e - 1]))) return -1;
    }
    File out(path, append);
    if (!out.f) return -2;
    std::string line;
    char num[32];
    for (int64_t e = 0; e < n_events;) {
        const int64_t row = ev_row[e];
        fixed16(time_us[row] * US_TO_S, num);
        if (format == EDENHIP_EVENTS_ASCII_V0) {
            line = num;
            for (; e < n_events && ev_row[e] == row; e++) {
                line += ' ';
                line += std::to_string(ids[ev_col[e]]);
            }
            line += '\n';
        } else {
            // ID_TIME prints the timestamp without its padding
            std::string t(num);
            const size_t b = t.find_first_not_of(' ');
            const size_t l = t.find_last_not_of(' ');
            const std::string trimmed = b == std::string::npos ? std::string() : t.substr(b, l - b + 1);
            line.clear();
            for (; e < n_events && ev_row[e] == row; e++) {
                const std::string id = std::to_string(ids[ev_col[e]]);
                line += format == EDENHIP_EVENTS_TIME_ID ? t + " " + id : id + " " + trimmed;
                line += '\n';
            }
        }
        if (std::fwrite(line.data(), 1, line.size(), out.f) != line.size()) return -2;
    }
    return std::ferror(out.f) ? -2 : 0;
}

}  // extern "C"
