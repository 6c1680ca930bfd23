#!/bin/bash
# Kernel profile of one full-size cornell_box job (1024x1024, 16+3x4 spp):
#   1. rocprofv3 --kernel-trace --stats  (per-kernel durations)
#   2. rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE  (HBM KiB per dispatch)
#   3. the same two counters on a plain 1 GiB device copy, to scale them to bytes
# and writes the summaries (markdown) into OUT/summary/.
# Usage: profile_kernels.sh [OUT]   (run from the repository root, GPU visible)
set -u
ROOT=$(cd "$(dirname "$0")/.." && pwd)
OUT=${1:-$ROOT/gpurun_out/profile}
mkdir -p "$OUT"
cd /tmp && export TMPDIR=/tmp
BENCH="python3 $ROOT/bench.py --steps 1 --warmup 0 --no-cpu-baseline --no-quality --no-served"
echo "profile: kernel trace"
timeout -k 10 300 rocprofv3 --kernel-trace --stats --output-format csv -d "$OUT/trace" -o run -- $BENCH > "$OUT/trace.log" 2>&1 || exit $?
for c in FETCH_SIZE WRITE_SIZE; do
  echo "profile: pmc $c"
  timeout -s KILL 120 rocprofv3 --pmc $c --output-format csv -d "$OUT/$c" -o run -- $BENCH > "$OUT/$c.log" 2>&1 || exit $?
  echo "profile: pmc $c on the calibration copy"
  timeout -s KILL 120 rocprofv3 --pmc $c --output-format csv -d "$OUT/cal_$c" -o run -- python3 $ROOT/bench.py --calibrate-copy > "$OUT/cal_$c.log" 2>&1 || exit $?
done
echo "profile: summary"
python3 "$ROOT/rpd-statmc-denoiser_amd/profile_summary.py" "$OUT" || exit $?
