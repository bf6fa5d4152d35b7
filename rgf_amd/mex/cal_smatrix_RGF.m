function [smatrix, channels] = cal_smatrix_RGF(syst, out, in)
%CAL_SMATRIX_RGF  S matrix of a 2D slab from librgf on an MI355X GPU.
%   Same arguments and outputs as cal_smatrix_RGF.m of chiaweihsu/RGF; the
%   solve runs in cal_smatrix_RGF_mex (librgf).  syst.precision may be
%   'fp64' (default) or 'fp32_refined'.
opts = struct();
if isfield(syst, 'precision'); opts.precision = syst.precision; end
smatrix = cal_smatrix_RGF_mex(syst, out, in, opts);
if nargout > 1
    channels = setup_channels(size(syst.epsilon, 1), syst.yBC, syst.k0dx, syst.epsilon_L, syst.epsilon_R);
end
end
