"""`python -m eden_amd nml <LEMS file> [rng_seed N] [gpu] [verbose] [chunk_rows N]`: EDEN's
command line (eden/Eden.cpp:10379-11968) for the GPU path.

Tokens as EDEN reads them (eden/Eden.cpp:10436-10546): `nml <file>` selects the model,
`rng_seed N` overrides the simulation seed, `verbose` prints more, `gpu` selects the GPU
path (the only one this package has); `chunk_rows N` sets the rows per GPU chunk. Unknown
tokens are skipped with a message, as EDEN does. Output files are written relative to the
working directory. At the end it prints EDEN's timing line
`Config: <s> Setup: <s> Run: <s>` and the memory line (eden/Eden.cpp:11951-11957). Exit
status 0; 1 when the model cannot be read or run; 2 when no model was selected.
"""
import resource
import sys
import time


def main(argv):
    t0 = time.time()
    path, seed, verbose, chunk = None, None, False, None
    i = 0
    while i < len(argv):
        a = argv[i]
        if a == "nml":
            if i == len(argv) - 1:
                print("cmdline: NeuroML filename missing")
                return 1
            path = argv[i + 1]
            i += 1
        elif a in ("rng_seed", "chunk_rows"):
            if i == len(argv) - 1:
                print(f"cmdline: {a} value missing")
                return 1
            try:
                v = int(argv[i + 1])
            except ValueError:
                print(f"cmdline: {a} value {argv[i + 1]} is not an integer")
                return 1
            if a == "rng_seed":
                seed = v
            else:
                chunk = v
            i += 1
        elif a == "verbose":
            verbose = True
        elif a == "gpu":
            pass
        else:
            print(f'cmdline: skipping unknown token "{a}"')
        i += 1
    if path is None:
        print("error: NeuroML model not selected (select one with nml <file> in command line)")
        return 2
    from . import neuroml
    from .model import EdenHipError
    from .run_sim import run_lems
    config_s = time.time() - t0
    try:
        res = run_lems(path, rng_seed=seed, chunk_rows=chunk, verbose=verbose)
    except (neuroml.NeuroMLError, EdenHipError, OSError) as e:
        print(f"NeuroML model could not be created: {e}")
        return 1
    print("Config: %.3lf Setup: %.3lf Run: %.3lf " % (config_s + res["config_s"], res["setup_s"], res["run_s"]))
    peak = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss * 1024
    try:
        with open("/proc/self/statm") as f:
            now = int(f.read().split()[1]) * resource.getpagesize()
    except OSError:
        now = -1
    print("Peak: %d Now: %d Heap: %d" % (peak, now, -1))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
