#!/bin/sh
# Builds the reference's own CORSIKA IACT reader (its src/bernlohr/ sources,
# compiled where they lie under /root/reference) with the dump driver
# oracle/corsika_dump.c into oracle/_ref/corsika_dump. Test infrastructure
# only; skipped when the reference tree is absent.
set -e
cd "$(dirname "$0")"
REF=/root/reference
[ -d "$REF/src/bernlohr" ] || { echo "reference tree absent: oracle/_ref not built"; exit 0; }
mkdir -p _ref
gcc -O1 -w -I"$REF/include/bernlohr" -o _ref/corsika_dump corsika_dump.c \
    "$REF/src/bernlohr/eventio.c" "$REF/src/bernlohr/io_simtel.c" "$REF/src/bernlohr/fileopen.c" \
    "$REF/src/bernlohr/warning.c" "$REF/src/bernlohr/straux.c" -lm
