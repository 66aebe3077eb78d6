"""The text EDEN's writers produce, built from the reference printer's own arithmetic (not
from the writer under test), for the byte-exact output tests.

The engine loop builds FixedWidthNumberPrinter(column_width = 16, delimiter '\\t', delimiter
length 0) (eden/Eden.cpp:10570-10571). Its constructor (eden/Eden.cpp:154-163) sets
number_size = column_size - delimiter_size and digits = column_size - 3 - 5 and formats
"%+<number_size>.<digits>g"; write() (eden/Eden.cpp:165-173) prints with snprintf bounded to
number_size + 1 bytes and appends delimiter_size delimiter characters. Trajectory lines are
the timestamp, then ' ' + number per column (eden/Eden.cpp:10856-10870); event lines TIME_ID
"<time> <id>", ID_TIME "<id> <trimmed time>", ascii_v0 "<time> <id> <id> ..."
(eden/Eden.cpp:10897-10927); '\\n' after every line (eden/Eden.cpp:10633-10640). Timestamps
are engine microseconds times 1e-6 (GetTimestampAscii, eden/Eden.cpp:184-189). Python's
%-formatting and C's printf both print the correctly rounded decimal.
"""
import numpy as np

COLUMN_SIZE, DELIM_CHAR, DELIM_SIZE = 16, "\t", 0
NUMBER_SIZE = COLUMN_SIZE - DELIM_SIZE
DIGITS = COLUMN_SIZE - 3 - 5
PRINTER_FORMAT = "%%+%d.%dg" % (NUMBER_SIZE, DIGITS)


def fx(v):
    """FixedWidthNumberPrinter::get(v): snprintf(buf, number_size + 1, format, v), then the
    delimiter characters; the string is cut at its terminator."""
    return (PRINTER_FORMAT % v)[:NUMBER_SIZE] + DELIM_CHAR * DELIM_SIZE


def trajectory_text(t_us, trace, scale=1e-3):
    """`scale`: one factor for every column, or one per column."""
    rows, cols = trace.shape
    sc = [float(x) for x in np.broadcast_to(np.asarray(scale, np.float64), (cols,))]
    return "".join(fx(t_us[r] * 1e-6) + "".join(" " + fx(float(trace[r, c]) * sc[c]) for c in range(cols)) + "\n"
                   for r in range(rows))


def events_text(t_us, rows, cols, ids, fmt):
    lines = []
    if fmt == "ascii_v0":
        for r in sorted(set(int(x) for x in rows)):
            lines.append(fx(t_us[r] * 1e-6) + "".join(" %d" % ids[c] for rr, c in zip(rows, cols) if rr == r))
    else:
        for r, c in zip(rows, cols):
            t = fx(t_us[r] * 1e-6)
            lines.append(t + " %d" % ids[c] if fmt == "TIME_ID" else "%d " % ids[c] + t.strip(" "))
    return "".join(line + "\n" for line in lines)


def assert_same_text(got, want):
    """Equal texts, else a failure naming the first differing line (pytest's own diff of two
    large strings takes minutes)."""
    if got == want:
        return
    g, w = got.splitlines(), want.splitlines()
    for i, (a, b) in enumerate(zip(g, w)):
        if a != b:
            raise AssertionError(f"line {i} differs:\n got  {a!r}\n want {b!r}")
    raise AssertionError(f"{len(g)} lines against {len(w)} (one is a prefix of the other)")
