#!/bin/sh
# Regenerates the golden fixtures (detect, packed, show, transforms) from the
# reference sources.
# Needs /root/reference (this container only); the GPU box never runs it.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REPO=$(cd "$HERE/../.." && pwd)
REF=${REF:-/root/reference}
make -C "$REPO/oracle" ref REF="$REF"
mkdir -p "$REPO/oracle/_ref"
gcc -O2 -std=gnu11 -ffp-contract=off -fno-fast-math -fopenmp \
    -I"$REF/src" -I"$REPO/vid.stab_amd/csrc/synth" \
    -o "$REPO/oracle/_ref/make_golden" "$HERE/make_golden.c" \
    "$REPO/vid.stab_amd/csrc/synth/vsynth.c" \
    -L"$REPO/oracle/_ref" -Wl,-rpath,"$REPO/oracle/_ref" -lvsref -lm
gcc -O2 -std=gnu11 -ffp-contract=off -fno-fast-math \
    -I"$REF/src" -o "$REPO/oracle/_ref/make_golden_l2t" "$HERE/make_golden_l2t.c" \
    -L"$REPO/oracle/_ref" -Wl,-rpath,"$REPO/oracle/_ref" -lvsref -lm
# make_golden.sh [planar|packed|show [case index]|transforms|l2t|hashes]: one fixture set, or all
# (hashes, the whole workload clips, is slow: run it on its own)
case "$1" in
  hashes) "$REPO/oracle/_ref/make_golden" "$HERE/clip_hashes.json" hashes ${THREADS:-8} ;;
  planar) "$REPO/oracle/_ref/make_golden" "$HERE/detect_golden.json" planar $2 ;;
  packed) "$REPO/oracle/_ref/make_golden" "$HERE/packed_golden.json" packed $2 ;;
  show) "$REPO/oracle/_ref/make_golden" "$HERE/show_golden.json" show $2 ;;
  transforms) "$REPO/oracle/_ref/make_golden" "$HERE/transforms_golden.json" transforms ;;
  l2t) "$REPO/oracle/_ref/make_golden_l2t" "$HERE/l2t_golden.json" ;;
  *) "$REPO/oracle/_ref/make_golden" "$HERE/detect_golden.json" planar
     "$REPO/oracle/_ref/make_golden" "$HERE/packed_golden.json" packed
     "$REPO/oracle/_ref/make_golden" "$HERE/show_golden.json" show
     "$REPO/oracle/_ref/make_golden" "$HERE/transforms_golden.json" transforms
     "$REPO/oracle/_ref/make_golden_l2t" "$HERE/l2t_golden.json" ;;
esac
