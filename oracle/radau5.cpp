// placeholder translation unit; the Radau IIA integrator is added separately
