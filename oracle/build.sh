#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/liboracle.so.
# -ffp-contract=off: no FMA contraction, so the restatement's rounding is the
# plain IEEE sequence of the written expressions.
set -e
cd "$(dirname "$0")"
g++ -O2 -std=c++17 -ffp-contract=off -fPIC -shared -Wall -o liboracle.so rb_oracle.cpp -lpthread
