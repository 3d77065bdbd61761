"""The device arena's allocator (parafrost_amd/csrc/pfs_arena.cpp) on the
host: tests/arena/arena_host.cpp builds it with the device runtime calls
stubbed by host memory (each stream's copies queued and run, when a stream
is waited on, interleaved at random with the other streams' queued copies,
so compaction that lets one context's copy overwrite what another's has not
read yet fails) and drives it with random allocations, releases,
in-place growth, releases on idle streams, drains of a stream's pending
releases and compactions across four clients in small fixed-size
arenas (256 KiB to 2 MiB), checking after every operation that every block's slot still
points at its own content, that no blocks overlap and that the per-client
accounting adds up; built with AddressSanitizer and UBSan."""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_arena_allocator_randomized():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "arena_host")
        subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover",
                        "-o", exe, os.path.join(HERE, "arena", "arena_host.cpp")], check=True)
        r = subprocess.run([exe, "12", "2000"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    # the randomized sizes make room run out: compaction and the failure path both run
    assert out["compactions"] > 1000 and out["moved_bytes"] > 0 and out["alloc_fails"] > 0 and out["grows"] > 0
