"""ffip_amd: MI355X (gfx950) FDTD time-stepping engine behind ffip's Python API."""
