#!/bin/bash
# Builds igap4_amd/libigap_gpu.so for gfx950 from igap4_amd/csrc.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
"$HIPCC" -O3 -std=c++17 --offload-arch=gfx950 -fPIC -shared \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    ${IGAP_EXTRA_FLAGS} \
    "$HERE/csrc/igap_kernels.hip" "$HERE/csrc/igap_solver.hip" "$HERE/csrc/igap_ws.hip" \
    "$HERE/csrc/igap_capi.cpp" "$HERE/csrc/igap_ws_capi.cpp" \
    -lrocsparse -lrocblas -lrocsolver \
    -o "$HERE/libigap_gpu.so.tmp"
mv "$HERE/libigap_gpu.so.tmp" "$HERE/libigap_gpu.so"
# the plain-C host of the ABI (tests/c_host), linked against the library in place
gcc -O2 -std=c99 -Wall -Wextra -I "$HERE/../include" "$HERE/../tests/c_host/igap_c_host.c" \
    -L "$HERE" -ligap_gpu -Wl,-rpath,'$ORIGIN/../../igap4_amd' -Wl,-rpath-link,/opt/rocm/lib -lm \
    -o "$HERE/../tests/c_host/igap_c_host"
