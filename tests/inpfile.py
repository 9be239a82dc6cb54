"""Writes DENISE .inp parameter files (the reference's positional format:
one "key = value" line per parameter, in the order src/read_par.c reads
them, '#' comment lines ignored) for the CLI tests.  Keys need at least two
characters: the reader drops a line's first character before the key."""

ORDER = [
    "MODE", "PHYSICS", "NPROCX", "NPROCY", "FD_ORDER", "max_relative_error", "NX", "NY", "DH", "TIME", "DT",
    "QUELLART", "SIGNAL_FILE", "TS", "SRCREC", "SOURCE_FILE", "RUN_MULTIPLE_SHOTS", "FC_SPIKE_1", "FC_SPIKE_2",
    "ORDER_SPIKE", "WRITE_STF", "READMOD", "MFILE", "WRITEMOD", "relaxation_mechanisms_(L)", "FL", "TAU", "FREE_SURF", "FW", "DAMPING",
    "FPML", "npower", "k_max_PML", "BOUNDARY", "SNAP", "SNAP_SHOT", "TSNAP1", "TSNAP2", "TSNAPINC", "IDX", "IDY",
    "SNAP_FORMAT", "SNAP_FILE", "SEISMO", "READREC", "REC_FILE", "REFREC", "N_STREAMER", "REC_INCR_X",
    "REC_INCR_Y", "NDT", "SEIS_FORMAT", "SEIS_FILE_VX", "SEIS_FILE_VY", "SEIS_FILE_CURL", "SEIS_FILE_DIV",
    "SEIS_FILE_P", "LOG_FILE", "LOG", "ITERMAX", "JACOBIAN", "DATA_DIR", "TAPER", "TAPERLENGTH", "GRADT",
    "INVMAT1", "GRAD_FORM", "QUELLTYPB", "TESTSHOT", "SWS_TAPER_GRAD_VERT", "SWS_TAPER_GRAD_HOR",
    "EXP_TAPER_GRAD_HOR", "SWS_TAPER_GRAD_SOURCES", "SWS_TAPER_CIRCULAR_PER_SHOT", "SRTSHAPE", "SRTRADIUS",
    "FILTSIZE", "SWS_TAPER_FILE", "TFILE", "INV_MOD_OUT", "INV_MODELFILE", "VPUPPERLIM", "VPLOWERLIM",
    "VSUPPERLIM", "VSLOWERLIM", "RHOUPPERLIM", "RHOLOWERLIM", "QSUPPERLIM", "QSLOWERLIM", "GRAD_METHOD",
    "PCG_BETA", "NLBFGS", "MODEL_FILTER", "FILT_SIZE", "DTINV", "EPS_SCALE", "STEPMAX", "SCALEFAC", "TRKILL",
    "TRKILL_FILE", "PICKS_FILE", "MISFIT_LOG_FILE", "MIN_ITER", "GRAD_FILTER", "FILT_SIZE_GRAD", "TIMELAPSE",
    "DATA_DIR_T0", "RTMOD", "GRAVITY", "NGRAVB", "NZGRAV", "GRAV_TYPE", "BACK_DENSITY", "DFILE", "RTM_SHOT",
]
assert len(ORDER) == 115

DEFAULTS = dict(
    MODE=0, PHYSICS=1, NPROCX=1, NPROCY=1, FD_ORDER=8, max_relative_error=1, NX=100, NY=60, DH=10.0, TIME=0.3,
    DT=1e-3, QUELLART=1, SIGNAL_FILE="wavelet/w", TS=0.1, SRCREC=1, SOURCE_FILE="source.dat",
    RUN_MULTIPLE_SHOTS=1, FC_SPIKE_1=-5.0, FC_SPIKE_2=15.0, ORDER_SPIKE=5, WRITE_STF=0, READMOD=1,
    MFILE="model/m", WRITEMOD=0, FL="10.0", TAU=0.0, FREE_SURF=1, FW=10, DAMPING=1500.0, FPML=10.0,
    npower=4.0, k_max_PML=1.0, BOUNDARY=0, SNAP=0, SNAP_SHOT=1, TSNAP1=0.002, TSNAP2=0.1, TSNAPINC=0.05, IDX=1,
    IDY=1, SNAP_FORMAT=3, SNAP_FILE="snap/s", SEISMO=1, READREC=1, REC_FILE="receiver", REFREC="0.0 , 0.0",
    N_STREAMER=0, REC_INCR_X=0, REC_INCR_Y=0, NDT=1, SEIS_FORMAT=1, SEIS_FILE_VX="su/out_x.su",
    SEIS_FILE_VY="su/out_y.su", SEIS_FILE_CURL="su/out_rot.su", SEIS_FILE_DIV="su/out_div.su",
    SEIS_FILE_P="su/out_p.su", LOG_FILE="log/l.log", LOG=1, ITERMAX=1, JACOBIAN="jac/j", DATA_DIR="obs/d",
    TAPER=0, TAPERLENGTH=4, GRADT="1, 2, 3, 4", INVMAT1=1, GRAD_FORM=1, QUELLTYPB=1, TESTSHOT="1, 100, 1",
    SWS_TAPER_GRAD_VERT=0, SWS_TAPER_GRAD_HOR=0, EXP_TAPER_GRAD_HOR=2.0, SWS_TAPER_GRAD_SOURCES=0,
    SWS_TAPER_CIRCULAR_PER_SHOT=0, SRTSHAPE=1, SRTRADIUS=5.0, FILTSIZE=1, SWS_TAPER_FILE=0, TFILE="taper/t",
    INV_MOD_OUT=0, INV_MODELFILE="model/inv", VPUPPERLIM=6000.0, VPLOWERLIM=0.0, VSUPPERLIM=4000.0,
    VSLOWERLIM=0.0, RHOUPPERLIM=3000.0, RHOLOWERLIM=1000.0, QSUPPERLIM=100.0, QSLOWERLIM=10.0, GRAD_METHOD=2,
    PCG_BETA=2, NLBFGS=20, MODEL_FILTER=0, FILT_SIZE=5, DTINV=2, EPS_SCALE=0.01, STEPMAX=6, SCALEFAC=2.0,
    TRKILL=0, TRKILL_FILE="tk.dat", PICKS_FILE="picks_", MISFIT_LOG_FILE="misfit.dat", MIN_ITER=0,
    GRAD_FILTER=0, FILT_SIZE_GRAD=10, TIMELAPSE=0, DATA_DIR_T0="obs0/d", RTMOD=0, GRAVITY=0, NGRAVB=0,
    NZGRAV=0, GRAV_TYPE=1, BACK_DENSITY=1, DFILE="d.rho", RTM_SHOT=0,
)

WORKFLOW_HEADER = ("PRO TIME_FILT FC_low FC_high ORDER TIME_WIN GAMMA TWIN- TWIN+ INV_VP_ITER INV_VS_ITER "
                   "INV_RHO_ITER INV_QS_ITER SPATFILTER WD_DAMP WD_DAMP1 EPRECOND LNORM ROWI STF_INV "
                   "OFFSETC_STF EPS_STF NORMALIZE OFFSET_MUTE OFFSETC SCALERHO SCALEQS ENV GAMMA_GRAV N_ORDER")


def write_inp(path, **overrides):
    vals = dict(DEFAULTS)
    vals["relaxation_mechanisms_(L)"] = overrides.pop("L", 0)
    vals.update(overrides)
    lines = ["#-------- parameter file (test) --------"]
    for k in ORDER:
        lines.append(f"# {k}")
        lines.append(f"{k} = {vals[k]}")
    lines.append("#")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_workflow(path, n_order=0, lnorm=2, time_filt=0, fc_start=0.0, fc_end=0.0, order=6, eprecond=0,
                   offset_mute=0, offsetc=10.0, spatfilter=0, wd_damp=0.5, wd_damp1=0.5, timewin=0, gamma=1.0, twlength_minus=0.0,
                   twlength_plus=0.0, inv_stf=0, offsetc_stf=-4.0, eps_stf=0.1, normalize=0):
    row = (f"0.01 {time_filt} {fc_start} {fc_end} {order} {timewin} {gamma} {twlength_minus} {twlength_plus} "
           f"0 0 0 0 {spatfilter} {wd_damp} {wd_damp1} "
           f"{eprecond} {lnorm} 0 {inv_stf} "
           f"{offsetc_stf} {eps_stf} {normalize} {offset_mute} {offsetc} 0.5 1.0 0 0 {n_order}")
    with open(path, "w") as f:
        f.write(WORKFLOW_HEADER + "\n" + row + "\n")
