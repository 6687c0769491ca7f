#!/bin/bash
# Builds and runs es23_16_values.f90 with amdflang (the compiler of the
# reference build, oracle/Makefile) to write es23_16_values.txt here.
set -e
cd "$(dirname "$0")"
tmp=$(mktemp -d)
/opt/rocm/bin/amdflang -O2 es23_16_values.f90 -o "$tmp/es"
(cd "$tmp" && ./es)
cp "$tmp/es23_16_values.txt" .
rm -rf "$tmp"
