function [smatrix, channels] = cal_smatrix_RGF(syst, out, in)
%CAL_SMATRIX_RGF  S matrix of a 2D slab from librgf on an MI355X GPU.
%   Same arguments and outputs as cal_smatrix_RGF.m of chiaweihsu/RGF; the
%   solve runs in cal_smatrix_RGF_mex (librgf).  syst.precision may be
%   'fp64' (default) or 'fp32_refined'.  channels comes from librgf too:
%   its fun_phi evaluates the transverse profiles in cal_smatrix_RGF_mex.
opts = struct();
if isfield(syst, 'precision'); opts.precision = syst.precision; end
smatrix = cal_smatrix_RGF_mex(syst, out, in, opts);
if nargout > 1
    channels = cal_smatrix_RGF_mex('channels', syst);
    % the profiles depend only on the transverse size and boundary condition
    lead = struct('epsilon', zeros(size(syst.epsilon, 1), 0), 'epsilon_L', 1, 'epsilon_R', 1, ...
                  'k0dx', 1, 'yBC', syst.yBC);
    channels.fun_phi = @(kydx) cal_smatrix_RGF_mex('phi', lead, kydx);
end
end
