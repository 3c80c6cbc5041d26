"""Autotunes fftset sizes on the current GPU and writes a tune table.

    python -m fftset_amd.tune --out fftset_amd/tune_mi355x.txt \\
        --cx 48,120,3840 --fo 2048,4096,61440

Each (modulation, size) is timed with fftset_autotune (candidate radix
chains x packings, or four-step splits) on a fresh set; the table is then
loaded by FFTSet(tune_file=...) or FFTSET_AMD_TUNE_FILE.
"""
import argparse
import json
import sys

import fftset_amd


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--cx", default="", help="COMPLEX sizes, comma separated")
    ap.add_argument("--fo", default="", help="FREQOFFSETREAL sizes, comma separated")
    a = ap.parse_args(argv)
    fs = fftset_amd.FFTSet()
    todo = [(fftset_amd.COMPLEX, int(v)) for v in a.cx.split(",") if v] + \
           [(fftset_amd.FREQ_OFFSET_REAL, int(v)) for v in a.fo.split(",") if v]
    for mod, L in todo:
        ms = fs.autotune(mod, L)
        f = fs.create_fft(mod, L)
        print(json.dumps({"mod": mod.split("_")[-1], "L": L, "ms_per_64MiB": round(ms, 5),
                          "radices": f.radices()}), flush=True)
    fs.tune_save(a.out)
    fs.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
