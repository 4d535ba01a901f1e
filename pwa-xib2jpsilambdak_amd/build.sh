#!/bin/bash
# Builds libpwa_amd.so (HIP kernels for gfx950 + host C++ + C ABI) in-tree.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/.." && pwd)"
OUT="$HERE/pwa_amd/libpwa_amd.so"
mkdir -p "$HERE/_build"
HIPFLAGS="-O3 -std=c++17 -fPIC --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops -I$ROOT/include -I$HERE/csrc"
hipcc $HIPFLAGS -c "$HERE/csrc/pwa_kernels.hip" -o "$HERE/_build/pwa_kernels.o"
hipcc $HIPFLAGS -x hip -c "$HERE/csrc/pwa_host.cpp" -o "$HERE/_build/pwa_host.o"
hipcc -O3 -std=c++17 -fPIC -I$ROOT/include -I$HERE/csrc -c "$HERE/csrc/pwa_couplings.cpp" -o "$HERE/_build/pwa_couplings.o"
hipcc -O3 -std=c++17 -fPIC -I$ROOT/include -I$HERE/csrc -c "$HERE/csrc/pwa_kinematics.cpp" -o "$HERE/_build/pwa_kinematics.o"
hipcc -shared -fPIC --offload-arch=gfx950 -o "$OUT.tmp" "$HERE/_build/pwa_kernels.o" "$HERE/_build/pwa_host.o" \
    "$HERE/_build/pwa_couplings.o" "$HERE/_build/pwa_kinematics.o"
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
# host-side regression check of the K-matrix adjoint (run by tests/test_kmatrix_adjoint_cpu.py)
hipcc -O2 -std=c++17 -I$ROOT/include -I$HERE/csrc "$ROOT/tests/cpp/km_adjoint_check.cpp" -o "$HERE/_build/km_adjoint_check"
