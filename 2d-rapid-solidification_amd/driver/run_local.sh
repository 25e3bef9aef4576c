#!/usr/bin/env bash
# run_local.sh -- local (no batch scheduler) launcher of one run, with the
# variables of the reference's run script (run.sh:3-30): the values it bakes
# into the executable with -D flags are passed to rsolid_run as run-time
# keys, so one build serves every run. Set any variable in the environment
# to override its default, e.g.
#
#   velocity=0.2 gradient=0.0236 total_time=20000 ./run_local.sh
#
# The run directory is ${path_input} (default ./<tag>); stdout goes to
# out_temp_<tag>.txt there and is appended to out.txt between rule lines,
# as the reference's job script does (run.sh:91-95). DRY_RUN=1 prints the
# command instead of running it. Extra arguments are passed on as further
# rsolid_run keys (precision=mixed, preset=config N=4096, state=1, ...).
set -euo pipefail

here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
exe="${RSOLID_RUN:-${here}/../bin/rsolid_run}"

Nx_all="${Nx_all:-$((12 * 64))}"
Ny_all="${Ny_all:-$((13 * 64))}"
material="${material:-Ti-Nb}"
name="${name:-ramped-slow}"
velocity="${velocity:-0.0842}"        # m/s, pulling velocity
gradient="${gradient:-0.0236}"        # K/nm, temperature gradient
total_time="${total_time:-120000}"    # ns
if_load="${if_load:-0}"
eps1="${eps1:-0.012}"                 # anisotropy of the interfacial free energy
epk1="${epk1:-0.1}"                   # kinetic anisotropy
run_time="${run_time:-4}"             # h, wall-time budget of the run
gpu_num=1

tag="${material}:${name}-${gpu_num}-${velocity}-${gradient}"
path_input="${path_input:-$(pwd)/${tag}}"
out_name="${path_input}/out_temp_${tag}.txt"

cmd=("${exe}" "Nx_all=${Nx_all}" "Ny_all=${Ny_all}" "total_time=${total_time}"
     "Vp=${velocity}" "GG=${gradient}" "eps1=${eps1}" "epk1=${epk1}"
     "path_input=${path_input}" "if_load=${if_load}" "run_time=${run_time}" "$@")

if [[ "${DRY_RUN:-0}" == "1" ]]; then
    printf '%q ' "${cmd[@]}"
    printf '\n'
    exit 0
fi

mkdir -p "${path_input}/data"
rc=0
"${cmd[@]}" > "${out_name}" || rc=$?
{
    printf '\n########################################\n'
    cat "${out_name}"
    printf '\n########################################\n'
} >> "${path_input}/out.txt"
exit "${rc}"
