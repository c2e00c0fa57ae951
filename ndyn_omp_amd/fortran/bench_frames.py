"""Wall time of the deck driver's output path: the same deck with its frames formatted by
write_model ('FrameWriter fortran') and by the library's writer thread ('FrameWriter engine'),
and whether the two files are identical.

    python ndyn_omp_amd/fortran/bench_frames.py [cells]

Cu Morse FCC, 4*cells^3 atoms (default 32: 131,072 atoms), 0.05-time-unit steps from a
Maxwell-Boltzmann start, skin trigger; 400 steps with thermo every 20 steps and frames every
0 (only the first and last), 100 or 20 steps in the dump and xyz formats.  Prints the wall
time of the whole driver run and the 'Elapsed time for main dynamics' line (the time loop
with its frames).
"""
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from ndyn_omp_amd import workloads as W  # noqa: E402

DRIVER = os.path.join(ROOT, "ndyn_omp_amd", "bin", "ndyn_gpu")


def write_model(path, n, a=3.615):
    pos, box = W.fcc(n, n, n, a)
    with open(path, "w") as f:
        f.write("ITEM: TIMESTEP\n0\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n" % len(pos))
        for L in box:
            f.write(f"0 {L:g}\n")
        f.write("ITEM: ATOMS id type xs ys zs\n")
        for k, p in enumerate(pos + 0.5):
            f.write(f"{k + 1} 1 {p[0]:.7f} {p[1]:.7f} {p[2]:.7f}\n")
    return len(pos)


def main():
    cells = int(sys.argv[1]) if len(sys.argv) > 1 else 32
    d = tempfile.mkdtemp()
    npart = write_model(os.path.join(d, "model.dump"), cells)
    print(f"atoms {npart}", flush=True)
    for conf, fmt in ((0, "dump"), (100, "dump"), (20, "dump"), (100, "xyz"), (20, "xyz")):
        out = {}
        for writer in ("fortran", "engine", "fortran", "engine"):
            deck = f"""START
Model model.dump dump
Seed 1234567890
Skin 0.5
Types 1
 Cu 63.546
Pairs 1
 Cu Cu sym morse 6.0 0.3429 1.359 2.866
Velocity 0.03877798635
OutFreq 20
ListTrigger skin
ConfOutFreq {conf} {fmt}
FrameWriter {writer}
Run 400 0.05
END
"""
            target = os.path.join(d, "out." + fmt)
            if os.path.exists(target):
                os.remove(target)
            t0 = time.perf_counter()
            p = subprocess.run([DRIVER], input=deck, cwd=d, capture_output=True, text=True,
                               timeout=300)
            t1 = time.perf_counter()
            if p.returncode != 0:
                sys.exit(f"driver failed:\n{p.stdout[-2000:]}\n{p.stderr[-2000:]}")
            data = open(target, "rb").read()
            out.setdefault(writer, data)
            loop = [l.split()[1] for l in p.stdout.splitlines() if l.strip().endswith("seconds")]
            print(f"ConfOutFreq {conf:3d} {fmt:4s} {writer:7s}: wall {t1 - t0:6.3f} s, "
                  f"time loop {float(loop[-1]):6.3f} s, file {len(data) / 1e6:6.1f} MB", flush=True)
        print(f"  identical files: {out['engine'] == out['fortran']}", flush=True)


if __name__ == "__main__":
    main()
