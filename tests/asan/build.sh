#!/bin/bash
# Host-side AddressSanitizer build of libgprx_hip's sources and of the driver
# fit_asan.cpp (batched L-BFGS fits with and without a jitter policy, a mixed
# fit, and uneven batches in both precisions).  Only host code is
# instrumented (-Xarch_host); the device code is the library's own.  Run the
# result on a GPU with ASAN_OPTIONS=detect_leaks=0:halt_on_error=1.
set -e
cd "$(dirname "$0")"
mkdir -p obj
for f in ../../gprx_amd/csrc/*.hip; do
  b=$(basename "$f" .hip)
  /opt/rocm/bin/hipcc -std=c++17 -O1 -g --offload-arch=gfx950 -fPIC \
    -Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer \
    -Xclang -target-feature -Xclang -packed-fp32-ops -c "$f" -o "obj/$b.o" &
done
wait
g++ -std=c++17 -O1 -g -fsanitize=address -fno-omit-frame-pointer -c fit_asan.cpp -o obj/fit_asan.o
g++ -fsanitize=address -o fit_asan obj/*.o -L/opt/rocm/lib -lamdhip64 -Wl,-rpath,/opt/rocm/lib
rm -rf obj
