function d = xgpudesc(p, k)
%   d = XGPUDESC(p) maps a prepared xSPDE structure (after xpreferences) whose
%   drift, linear term, initial state and observables are built-in GPU forms
%   named in p.gpu onto the descriptor of the MI355X library (xsh_desc in
%   include/xspde_hip.h), for xsh_mex('create', d).
%   d = XGPUDESC(p, k) is the descriptor of field component k of a project
%   with p.fields > 1: the built-in forms act on each component by itself, so
%   each component is its own descriptor, with row k of p.boundaries{1,d} and
%   p.boundval{1,d}, row k of the fields x 2 cell p.boundfun returns, the
%   overrides p.gpu.fields{k}, p.noises/p.fields noise components, the
%   observables o with p.gpu.ofield(o) == k and the streams of instance k-1.
%   p.gpu fields (the names of xspde_matlab_amd/xsim.py make_desc):
%     precision 'fp32' | 'fp64'
%     deriv     'zero' | 'kubo' (c) | 'gp' (g, u, s)
%     linear    struct with l0, l1 (1 or 2 values), l2 (1 or 2 values)
%     initial   'const' (c0) | 'sech' (amp, width, eps, center)
%               | 'gauss' (amp, width, width_y)
%               | 'phase' (amp) | 'coherent' (alpha)
%     observe   cell of 'a','abs2','a_ave','abs2_ave','abs2_int','abs4_ave',
%               'corr','lowk','abs2k'
%   p.boundfun and the k-space filters p.nfilter, p.rfilter are tabulated
%   here (boundtab, nfilter, rfilter); a boundfun that reads the fields
%   raises, and the project then stays on the CPU xensemble.
%   Not executed in the repository's tests (MATLAB is not part of its image).

if nargin < 2, k = 1; end
g = p.gpu;
F = p.fields{1};
if isfield(g,'fields') && numel(g.fields) >= k  % per-component overrides
  o = g.fields{k};
  fn = fieldnames(o);
  for i = 1:numel(fn), g.(fn{i}) = o.(fn{i}); end
end
obs = g.observe;
of = ones(1,numel(obs));
if isfield(g,'ofield'), of = g.ofield; end
mine = find(of == k);
nd = p.dimensions;
pts = p.points{1};
d.precision = double(isfield(g,'precision') && any(strcmpi(g.precision,{'fp64','double'})));
d.dimensions = nd;
d.points = pts(1);
d.nx = 1; d.ny = 1;
if nd > 1, d.nx = pts(2); end
if nd > 2, d.ny = pts(3); end
d.steps = p.steps(1);
d.checks = p.checks(1);
d.method = double(~strcmp(func2str(p.method),'RK4'));
d.iterations = p.iterations;
d.ensembles = p.ensembles;
d.seed = p.seed;
d.noises = field0(p,'noises')/F;
d.inrandoms = field0(p,'inrandoms')/F;
d.urandoms = field0(p,'urandoms')/F;
d.inst = k-1;
b = zeros(2,2);
for kd = 1:nd-1
  bd = p.boundaries{1,kd+1};
  b(kd,:) = bd(min(k,size(bd,1)),:);
end
d.boundaries = b;
d.ranges = zeros(1,3);  d.ranges(1:nd) = p.ranges(1:nd);
d.origins = zeros(1,3); d.origins(1:nd) = p.origins(1:nd);
names = {'zero','kubo','gp','matrix'};
d.deriv = find(strcmp(names, g.deriv)) - 1;
if strcmp(g.deriv,'matrix')                      % coupled SDE components: one descriptor
  % da_i/dt = sum_j M(i,j) a_j + g(i) |a_i|^2 a_i + sum_k S(i,k) w_k
  d.fields = F; d.inst = 0;
  d.noises = field0(p,'noises'); d.inrandoms = field0(p,'inrandoms'); d.urandoms = field0(p,'urandoms');
  d.fmat = complex(g.M);
  d.fnl = complex(getdef(g,'g',zeros(1,F)) + zeros(1,F));
  d.fnoise = getdef(g,'S',zeros(F,max(1,d.noises)));
  d.fc0 = complex(getdef(g,'c0',ones(1,F)) + zeros(1,F));
  d.fl0 = complex(zeros(1,F));
  if isfield(g,'linear'), d.fl0 = complex(getdef(g.linear,'l0',0) + zeros(1,F)); end
  mine = 1:numel(obs);
  d.obs_field = of - 1;
end
dp = zeros(1,6);
switch g.deriv
  case 'kubo', dp(1) = g.c;
  case 'gp',   dp(1:5) = [real(g.g), imag(g.g), real(g.u), imag(g.u), g.s];
end
d.deriv_par = dp;
d.has_linear = double(isfield(g,'linear'));
if d.has_linear
  L = g.linear;
  l0 = getdef(L,'l0',0); d.lin_l0 = complex(l0(1));  % (coupled components: fl0 per component)
  d.lin_l1 = complex(two(getdef(L,'l1',0)));
  d.lin_l2 = complex(two(getdef(L,'l2',0)));
end
inames = {'const','sech','gauss','phase','coherent'};
d.initial = find(strcmp(inames, g.initial)) - 1;
ip = zeros(1,6);
switch g.initial
  case 'const',    c = getdef(g,'c0',1); ip(1:2) = [real(c(1)), imag(c(1))];
  case 'sech',     ip(1:4) = [getdef(g,'amp',1), getdef(g,'width',1), getdef(g,'eps',0), getdef(g,'center',0)];
  case 'gauss',    ip(1:3) = [getdef(g,'amp',1), getdef(g,'width',1), getdef(g,'width_y',0)];
  case 'phase',    ip(1) = getdef(g,'amp',1);
  case 'coherent', a = getdef(g,'alpha',1); ip(1:2) = [real(a), imag(a)];
end
d.init_par = ip;
onames = {'a','abs2','a_ave','abs2_ave','abs2_int','abs4_ave','corr','lowk','abs2k'};
if isempty(mine)                                 % a component nobody observes
  d.observe = find(strcmp(onames,'a_ave')) - 1;
  d.tspec = 0;
else
  d.observe = zeros(1,numel(mine));
  d.tspec = zeros(1,numel(mine));
  for i = 1:numel(mine)
    d.observe(i) = find(strcmp(onames, obs{mine(i)})) - 1;
    % time-domain spectra: p.transforms{o}(1) (xpath.m:90-106), SDEs
    if isfield(p,'transforms') && mine(i) <= numel(p.transforms)
      d.tspec(i) = double(p.transforms{mine(i)}(1) ~= 0);
    end
  end
end
d.raw = double(isfield(p,'rawdata') && p.rawdata);
bv = zeros(2,2);
if isfield(p,'boundval')
  for kd = 1:nd-1
    v = p.boundval{1,kd+1};
    if ~isempty(v), r = min(k,size(v,1)); bv(kd,:) = [v{r,1}(1), v{r,2}(1)]; end
  end
end
d.boundval = complex(bv);
% p.boundfun at every half fine step (xpatch.m:37, xsetbound.m:39 call it at p.t)
if isfield(p,'boundfun') && ~isequal(func2str(p.boundfun),'xboundfun')
  nb = 1; if nd == 3, nb = max(d.nx, d.ny); end
  calls = d.points*d.steps - 1;
  levels = 2*calls*(1+d.checks) + 1;
  dtr = p.dt/(d.steps*(1+d.checks));
  tab = complex(zeros(nb,2,2,levels));
  for m = 1:levels
    p.t = p.origins(1) + (m-1)*dtr/2;
    for kd = 1:nd-1
      v = p.boundfun({[]},1,kd+1,p);
      r = min(k,size(v,1));                      % row k of a fields x 2 cell
      for side = 1:2
        line = v{r,side}(:);
        tab(1:numel(line),side,kd,m) = line;
      end
    end
  end
  d.boundtab = tab;
end
% k-space filters as multipliers at every k site in FFT order (xnoise.m:42, xrandom.m:26)
if field0(p,'knoises') > 0
  d.nfilter = kfilter(p, p.nfilter{1}, field0(p,'knoises')/F);
end
if field0(p,'krandoms') > 0
  d.rfilter = kfilter(p, p.rfilter{1}, field0(p,'krandoms')/F);
end
end

function v = field0(p, name)
v = 0;
if isfield(p, name) && ~isempty(p.(name)), v = p.(name)(1); end
end

function v = getdef(s, name, dflt)
v = dflt;
if isfield(s, name), v = s.(name); end
end

function v = two(x)
v = [x(1), x(end)];
end

function F = kfilter(p, f, count)
% the filter applied to all-ones arrays of the cell shape: [nx*ny, count]
sz = [count, p.points{1}(2:end), 1];
F = f(ones(sz), p);
F = reshape(F, count, []).';
end
