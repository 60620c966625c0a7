"""bench.py's contract pieces that need no GPU: the committed rocprof
summaries it reads its HBM traffic from carry the kernel names it labels its
line with, a request for more than one GPU is refused with a message, and the
warm-up fills the ring and the timed region holds whole time-blocked passes."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_profiled_traffic_matches_bench_labels():
    sys.path.insert(0, ROOT)
    import bench
    # the N = 65536 line's kernel (the time-blocked method's pass), and the
    # per-step method's (the default pipelined source loop)
    tb = bench.profiled_traffic(bench.kernel_name({"dynamic_items": 0, "pipelined": 1, "block_steps": 16}, 1, 2))
    assert tb is not None and 5e8 < tb < 1e10
    t32 = bench.profiled_traffic(bench.kernel_name({"dynamic_items": 0, "pipelined": 1}, 1, 0))
    assert t32 is not None and 5e8 < t32 < 1e10
    ttab = bench.profiled_traffic("pair_table_f64<4, false>", "pair_table_n16384.md")
    assert ttab is not None and ttab > 1e10
    assert bench.profiled_traffic("no_such_kernel<1>") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


class _FakeContext:
    """Records the step counts bench.timed_steps asks a context to run."""

    def __init__(self, block_steps):
        self.block_steps = block_steps
        self.runs = []

    def info(self):
        return {"block_steps": self.block_steps, "max_delay_bound": 40}

    def run(self, n):
        self.runs.append(n)

    def sync(self):
        pass

    def set_timing(self, on):
        pass

    def pair_kernel_time(self, reset=False):
        return 0.0, 0

    def short_kernel_time(self, reset=False):
        return 0.0, 0


def test_timed_region_holds_whole_passes():
    """The warm-up fills the ring (largest delay 40 + order 4 + 2 = 46 steps,
    rounded up to 48) whatever is asked for, and with K steps per pass it ends
    on a pass boundary and the timed steps are a multiple of K: 20 requested
    steps with 16-step passes time 32 (two whole passes), never one pass for
    20 steps."""
    sys.path.insert(0, ROOT)
    import bench
    for k, warm, steps, want_warm, want_steps in ((16, 5, 20, 48, 32), (8, 5, 20, 48, 24),
                                                    (16, 304, 192, 304, 192), (1, 5, 20, 48, 20),
                                                    (16, -1, 16, 48, 16), (8, 100, 20, 104, 24)):
        ctx = _FakeContext(k)
        out = bench.timed_steps(ctx, warm, steps)
        assert out["warmup"] == want_warm and out["steps"] == want_steps
        assert ctx.runs == [want_warm, want_steps, 16, want_steps]
