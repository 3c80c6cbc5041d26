"""Autotunes fftset sizes on the current GPU and writes a tune table.

    python -m fftset_amd.tune --out fftset_amd/tune_mi355x.txt \\
        --cx 48,120,3840:fwd+inv --fo 2048:conv+grad,4096,61440 \\
        [--cx-ops fwd] [--fo-ops conv]

Each (modulation, size) is timed with fftset_autotune (candidate radix
chains x packings, or four-step splits) on the operations it will run:
fwd, inv, kern, conv, grad, joined by '+' after the size, else the
modulation's --cx-ops / --fo-ops default. The table is then loaded by
FFTSet(tune_file=...) or FFTSET_AMD_TUNE_FILE.
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
    ap.add_argument("--cx-ops", default="fwd+conv", help="default operations for COMPLEX sizes")
    ap.add_argument("--fo-ops", default="fwd+conv", help="default operations for FREQOFFSETREAL sizes")
    a = ap.parse_args(argv)
    names = {"fwd": fftset_amd.OP_FORWARD, "inv": fftset_amd.OP_INVERSE, "kern": fftset_amd.OP_GET_KERNEL,
             "conv": fftset_amd.OP_CONV, "grad": fftset_amd.OP_CONV_GRAD_INPUT}

    def entries(mod, spec, default):
        out = []
        for item in spec.split(","):
            if not item:
                continue
            size, _, ops = item.partition(":")
            out.append((mod, int(size), [names[o] for o in (ops or default).split("+")]))
        return out

    fs = fftset_amd.FFTSet()
    todo = entries(fftset_amd.COMPLEX, a.cx, a.cx_ops) + entries(fftset_amd.FREQ_OFFSET_REAL, a.fo, a.fo_ops)
    for mod, L, ops in todo:
        ms = fs.autotune(mod, L, ops)
        f = fs.create_fft(mod, L)
        print(json.dumps({"mod": mod.split("_")[-1], "L": L, "ms_per_64MiB": round(ms, 5),
                          "radices": f.radices()}), flush=True)
    fs.tune_save(a.out)
    fs.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
