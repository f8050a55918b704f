"""Synthetic turbulence for the full-size GPU tests, generated in a child
process by particle_transport_amd.fields (torch FFTs there when a GPU is
visible): the test process loads only libpt_hip.so's HIP runtime."""
import os
import subprocess
import sys

from particle_transport_amd import fields

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def generate(n_nodes, *, seed, dB, kmax, modes=False):
    """A list of 6-tuples (Bx, By, Bz, ux, uy, uz) of (n,n,n) float32 arrays in
    SI units: the total field, then Alfven/fast/slow with modes=True."""
    cmd = [sys.executable, "-m", "particle_transport_amd.fields", "--nodes", str(n_nodes), "--seed", str(seed),
           "--dB", str(dB), "--kmax", str(kmax)] + (["--modes"] if modes else [])
    proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, cwd=ROOT)
    out = fields.read_stream(proc.stdout)
    if proc.wait() != 0:
        raise RuntimeError("field generator failed")
    return out
