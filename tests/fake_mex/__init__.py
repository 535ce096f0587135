"""Test-only fake MATLAB MEX host for the gateway in voxhenry_amd/mex."""
