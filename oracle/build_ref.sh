#!/usr/bin/env bash
# build_ref.sh -- compile the reference CADMAS-SURF/3D solver from its own
# Fortran sources (read in place under /root/reference) into oracle/_ref/.
#
# TEST INFRASTRUCTURE ONLY: the resulting executable generates golden fixtures
# (tests/golden/make_golden.py) and serves as the in-container timing
# reference.  Nothing under oracle/_ref is linked into the product.
#
# Toolchain: AMD flang (ROCm LLVM) + the MPICH that ships in /opt/conda (F77
# bindings, libmpifort/libmpi).  No stand-in headers or libraries are used.
# Deviations from Source code/Makefile:
#   * -fpconstant (Intel-only) dropped; every hot-path literal is a D0 constant.
#   * vf_im2inp.o in SRCS is compiled from vf_im2inp.f.
#   * VF_V1CAL, VF_VPSOL, VF_F1CAL and VF_CDIV00 are renamed to *_orig_ in their
#     own objects so that oracle/ref_dump.c can record each phase's inputs and
#     outputs (active only when CADMAS_DUMP_DIR is set).
set -euo pipefail

REF="${CADMAS_REF_SRC:-/root/reference/Simulators/CADMAS-SURF-3D/Source code}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
OBJ="$OUT/obj"
FC="${FC:-/opt/rocm/lib/llvm/bin/flang}"
MPIINC=/opt/conda/include
MPILIB=/opt/conda/lib
FFLAGS="-O2"
JOBS="${JOBS:-8}"

if [ ! -d "$REF" ]; then
    echo "build_ref.sh: reference sources not found at $REF (skipping)" >&2
    exit 0
fi
mkdir -p "$OBJ"

# modules first (order matters for .mod files)
for f in mod_comm.f90 mod_fault.f mod_apara.f; do
    b="${f%.*}"
    if [ ! -f "$OBJ/$b.o" ] || [ "$REF/$f" -nt "$OBJ/$b.o" ]; then
        (cd "$OBJ" && "$FC" -c $FFLAGS -I"$REF" -I"$MPIINC" "$REF/$f" -o "$b.o")
    fi
done

# the source list of the reference's own Makefile
SRCS=$(sed -n '/^SRCS/,/^OBJS/p' "$REF/Makefile" | grep -o '[a-z0-9_]*\.[fo]\b' | sed 's/\.o$/.f/' | grep -E '^[a-z0-9_]+\.f$' | grep -v '^mod_' | sort -u)
printf '%s\n' $SRCS | xargs -P "$JOBS" -I{} bash -c '
    f="{}"; b="${f%.f}"
    if [ ! -f "'"$OBJ"'/$b.o" ] || [ "'"$REF"'/$f" -nt "'"$OBJ"'/$b.o" ]; then
        cd "'"$OBJ"'" && "'"$FC"'" -c '"$FFLAGS"' -I"'"$REF"'" -I"'"$MPIINC"'" "'"$REF"'/$f" -o "$b.o.tmp" && mv "$b.o.tmp" "$b.o"
    fi'

# rename the wrapped phase routines inside their defining objects
for r in v1cal vpsol f1cal cdiv00; do
    objcopy --redefine-sym "vf_${r}_=vf_${r}_orig_" "$OBJ/vf_${r}.o" "$OBJ/vf_${r}.wrapped.o"
done

gcc -O2 -c "$HERE/ref_dump.c" -o "$OBJ/ref_dump.o"

OBJS=""
for f in $SRCS; do
    b="${f%.f}"
    case "$b" in
        vf_v1cal|vf_vpsol|vf_f1cal|vf_cdiv00) OBJS="$OBJS $OBJ/$b.wrapped.o" ;;
        *) OBJS="$OBJS $OBJ/$b.o" ;;
    esac
done
"$FC" -o "$OUT/cadmas_ref.x" "$OBJ/mod_comm.o" "$OBJ/mod_fault.o" "$OBJ/mod_apara.o" $OBJS \
    "$OBJ/ref_dump.o" -L"$MPILIB" -lmpifort -lmpi -Wl,-rpath,"$MPILIB"
echo "built $OUT/cadmas_ref.x"

# The reference relinked with the gfx950 library in place of its own solve
# routines (link-level drop-in, no reference file changed): the product's
# vf_m1bcgs_ (cadmas_ref_m1bcgs_hip.x) or vf_vpsol_ (cadmas_ref_vpsol_hip.x)
# is what the reference's own calls bind to.  -rdynamic exports the
# reference's COMMON blocks and VF_A2ERR to the library.  Built only when the
# library exists; the tests run these on the GPU.
PKG="$(cd "$HERE/.." && pwd)/multiscale-and-multiphysics-integrated-simulator-for-tsunami_amd"
if [ -f "$PKG/libcadmas_hip.so" ]; then
    for v in m1bcgs vpsol; do
        case "$v" in
            m1bcgs) drop="vf_m1bcgs" ;;
            vpsol) drop="vf_vpsol vf_m1bcgs" ;;
        esac
        DOBJS=""
        for f in $SRCS; do
            b="${f%.f}"
            skip=0
            for x in $drop; do [ "$b" = "$x" ] && skip=1; done
            [ $skip = 1 ] || DOBJS="$DOBJS $OBJ/$b.o"
        done
        "$FC" -o "$OUT/cadmas_ref_${v}_hip.x" "$OBJ/mod_comm.o" "$OBJ/mod_fault.o" "$OBJ/mod_apara.o" \
            $DOBJS -L"$PKG" -lcadmas_hip -rdynamic \
            -Wl,-rpath,'$ORIGIN/../../multiscale-and-multiphysics-integrated-simulator-for-tsunami_amd' \
            -L"$MPILIB" -lmpifort -lmpi -Wl,-rpath,"$MPILIB"
        echo "built $OUT/cadmas_ref_${v}_hip.x"
    done
fi
