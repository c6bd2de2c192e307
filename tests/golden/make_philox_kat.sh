#!/bin/sh
# Regenerates tests/golden/philox_kat.json from rocRAND's MIT headers (host build).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
TMP=$(mktemp -d)
hipcc -O1 -std=c++17 -o "$TMP/philox_kat" "$HERE/philox_kat.cpp"
"$TMP/philox_kat" > "$HERE/philox_kat.json"
rm -rf "$TMP"
