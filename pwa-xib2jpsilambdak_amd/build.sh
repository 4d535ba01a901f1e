#!/bin/bash
# Full rebuild of libpwa_amd.so in-tree (see Makefile); extra arguments go to make.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
JOBS="${MAX_JOBS:-$(nproc)}"
[ "$JOBS" -gt 16 ] && JOBS=16
make -C "$HERE" -B -j"$JOBS" "$@"
