#!/bin/sh
# Builds the Node-API addon against the system Node headers and links it to
# the cloth library next to it (closim_amd/_build), found through $ORIGIN.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/../_build"
NODE_INC=${NODE_INC:-/usr/include/node}
mkdir -p "$OUT"
gcc -O2 -std=c11 -fPIC -shared -Wall -Wextra -Wno-unused-parameter -Wno-missing-field-initializers \
  -DNODE_GYP_MODULE_NAME=hip_solver_addon -I"$NODE_INC" -I"$HERE/../../include" \
  -o "$OUT/hip_solver_addon.node.tmp" "$HERE/hip_solver_addon.c" \
  -L"$OUT" -lclothhip -Wl,-rpath,'$ORIGIN'
mv "$OUT/hip_solver_addon.node.tmp" "$OUT/hip_solver_addon.node"
