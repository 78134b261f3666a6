function s = optifax_export(obj, file)
%OPTIFAX_EXPORT Write an OptiFax object as JSON for the MI355X engine.
%   s = optifax_export(obj, file) encodes an OpticalSim, Cavity, Optic,
%   NonlinearCrystal, Laser or SimWindow object (the objects the toolbox
%   stores with obj.store as .mat files) as a struct of plain values and
%   writes it to `file` with jsonencode.  optifax_amd.interchange reads it.
%
%   Values: numbers and arrays as they are; strings as strings; function
%   handles as {"function": func2str(h)}; surfaces, bulks and optics as
%   nested structs with a "class" field.  Only properties the engine uses
%   are written (no plotting state, no transient fields).
arguments
	obj
	file string = ""
end
s = encode(obj);
s.format = "optifax-export";
s.version = 1;
if strlength(file) > 0
	fid = fopen(file, "w");
	fwrite(fid, jsonencode(s, "PrettyPrint", true), "char");
	fclose(fid);
end
end

function v = val(x)
if isa(x, "function_handle")
	v = struct("function", func2str(x));
elseif isstring(x) || ischar(x)
	v = string(x);
elseif isempty(x)
	v = [];
else
	v = x;
end
end

function s = encode(obj)
switch class(obj)
	case "OpticalSim"
		s = struct("class", "OpticalSim", ...
			"SimWin", encode(obj.SimWin), "Source", encode(obj.Source), ...
			"System", encode(obj.System), "Delay", obj.Delay, ...
			"RoundTrips", obj.RoundTrips, "StepSize", val(obj.StepSize), ...
			"AdaptiveError", val(obj.AdaptiveError), "ProgressPlots", obj.ProgressPlots, ...
			"Precision", string(obj.Precision), "PumpRadiusXtalIn", obj.PumpRadiusXtalIn, ...
			"DetectorPosition", obj.DetectorPosition);
	case "SimWindow"
		s = struct("class", "SimWindow", "ReferenceWave", obj.ReferenceWave, ...
			"NumberOfPoints", obj.NumberOfPoints, "TemporalRange", obj.TemporalRange, ...
			"TimeOffset", val(obj.TimeOffset), "SpectralLimits", obj.SpectralLimits);
	case "Laser"
		s = struct("class", "Laser", "Wavelength", obj.Wavelength, "Waist", obj.Waist, ...
			"RepetitionRate", obj.RepetitionRate, "AveragePower", obj.AveragePower, ...
			"PulseDuration", obj.PulseDuration, "SourceString", val(obj.SourceString), ...
			"PhaseString", val(obj.PhaseString));
	case "Cavity"
		optics = {};
		names = obj.Optics.Properties.VariableNames;
		for ii = 1:width(obj.Optics)
			o = encode(obj.Optics.(ii));
			o.Name = string(names{ii});
			optics{end+1} = o; %#ok<AGROW>
		end
		s = struct("class", "Cavity", "Name", val(obj.Name), "OCPosition", obj.OCPosition);
		s.Optics = optics;
	case {"Optic", "NonlinearCrystal"}
		s = struct("class", class(obj), "Regime", string(obj.Regime), ...
			"S1", surface(obj.S1), "S2", surface(obj.S2), ...
			"Bulk", struct("Material", string(obj.Bulk.Material), ...
				"Length", val(obj.Bulk.Length), "Temperature", obj.Bulk.Temperature));
		if class(obj) == "NonlinearCrystal"
			s.GratingPeriod = val(obj.GratingPeriod);
			s.Uncertainty = val(obj.Uncertainty);
			s.DutyCycleOffset = val(obj.DutyCycleOffset);
			s.StepSize = obj.StepSize;
			s.Height = obj.Height;
			s.VerticalPosition = obj.VerticalPosition;
			s.WaistPosition = obj.WaistPosition;
			s.Chi2 = obj.Chi2;
		end
	otherwise
		error("optifax_export:class", "cannot export a %s", class(obj));
end
end

function s = surface(surf)
s = struct("Coating", val(surf.Coating), "IncidentAngle", surf.IncidentAngle, ...
	"GDD", val(surf.GDD));
end
