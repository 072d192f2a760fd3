#!/bin/bash
# Sweep over Matrix-Market files: one CSV row per file and dense width, as
# the reference's run.sh does.  Columns: the 8 reference algorithms
# (sequential/parallel reduction x row/element balanced x row/col-major
# dense operands), then the rocSPARSE comparator columns (4 for nv = 1,
# 10 otherwise).  Extra arguments before the files are passed to the
# harness (e.g. --native adds a column for the library's own kernel).
#
#   da-spmmul-kernel_amd/run.sh [--native] file1.mtx file2.mtx ...
#   LOG_FILE=out.csv da-spmmul-kernel_amd/run.sh ...
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
LOG_FILE="${LOG_FILE:-./log.csv}"
OPTS=()
while [ $# -gt 0 ] && [ "${1#--}" != "$1" ]; do OPTS+=("$1"); shift; done
HEADER="filename,KDense,SR_RB_RM,PR_RB_RM,SR_EB_RM,PR_EB_RM,SR_RB_CM,PR_RB_CM,SR_EB_CM,PR_EB_CM"
HEADER="$HEADER,rocsparse_default,rocsparse_1,rocsparse_2,rocsparse_3,rocsparse_4"
HEADER="$HEADER,rocsparse_5,rocsparse_6,rocsparse_7,rocsparse_8,rocsparse_9"
for o in "${OPTS[@]}"; do [ "$o" = "--native" ] && HEADER="$HEADER,native"; done
echo "$HEADER" > "$LOG_FILE"
status=0
for f in "$@"; do
    "$HERE/bin/spmvspmm" "$f" "${OPTS[@]}" | tee -a "$LOG_FILE"
    rc=${PIPESTATUS[0]}
    [ "$rc" -eq 0 ] || status=$rc
done
exit $status
