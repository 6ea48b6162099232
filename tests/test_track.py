"""CPU check of the addon's page write tracking (sciwasm/sparse_amd/node/
write_track.h): a write to a protected page is caught exactly once, lands,
and leaves the page writable; reads never fault; memmove and concurrent
writers from eight threads are caught; the owner's own writes are counted
by unprotect_range; a fault outside the region still ends the process with
SIGSEGV (the previous/default action).  The addon keeps a device mirror of
a matrix array valid exactly while the write counts of its pages stay
unchanged, so these are the properties its residency cache rests on."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_write_tracking_host(tmp_path):
    exe = str(tmp_path / "track_host")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-pthread", "-o", exe,
                           os.path.join(ROOT, "tests", "native", "track_host.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
