#!/bin/sh
# Builds the CPU checker of the library's index arithmetic into $1.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
g++ -O2 -std=c++17 -shared -fPIC -I"$HERE/../../3d-ffts-mpi_amd/csrc" "$HERE/host_check.cpp" -o "$1"
