function [smatrix, channels] = cal_smatrix_RGF_3D(syst, out, in)
%CAL_SMATRIX_RGF_3D  S matrix of a 3D slab from librgf on an MI355X GPU.
%   Same arguments and outputs as cal_smatrix_RGF_3D.m of chiaweihsu/RGF;
%   channels comes from librgf too: its fun_phi(kxdx, kydx) evaluates the
%   transverse profile in cal_smatrix_RGF_mex.
opts = struct();
if isfield(syst, 'precision'); opts.precision = syst.precision; end
smatrix = cal_smatrix_RGF_mex(syst, out, in, opts);
if nargout > 1
    channels = cal_smatrix_RGF_mex('channels', syst);
    lead = struct('epsilon', zeros(size(syst.epsilon, 1), size(syst.epsilon, 2), 0), 'epsilon_L', 1, ...
                  'epsilon_R', 1, 'k0dx', 1, 'xyBC', syst.xyBC);
    channels.fun_phi = @(kxdx, kydx) cal_smatrix_RGF_mex('phi', lead, kxdx, kydx);
end
end
