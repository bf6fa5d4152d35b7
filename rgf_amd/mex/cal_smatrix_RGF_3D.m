function [smatrix, channels] = cal_smatrix_RGF_3D(syst, out, in)
%CAL_SMATRIX_RGF_3D  S matrix of a 3D slab from librgf on an MI355X GPU.
%   Same arguments and outputs as cal_smatrix_RGF_3D.m of chiaweihsu/RGF.
opts = struct();
if isfield(syst, 'precision'); opts.precision = syst.precision; end
smatrix = cal_smatrix_RGF_mex(syst, out, in, opts);
if nargout > 1
    channels = setup_channels_3D(size(syst.epsilon, 1), size(syst.epsilon, 2), syst.xyBC, syst.k0dx, ...
                                 syst.epsilon_L, syst.epsilon_R);
end
end
